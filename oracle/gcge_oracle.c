/*
 * gcge_oracle.c -- CPU restatement of GCGE-1.0's default solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker; see gcge_oracle.h).
 *
 * Structure: a per-vector CSR backend (app/csr), the default multivector
 * operations built from it (src/gcge_ops.c), and the GCG algorithm
 * (src/gcge_eigsol.c, gcge_xpw.c, gcge_cg.c, gcge_orthogonal.c,
 * gcge_rayleighritz.c) with the workspace sizing of src/gcge_workspace.c and
 * the parameter defaults of src/gcge_para.c.  Multivectors are arrays of
 * column pointers and MultiVecSwap exchanges pointers, exactly as the
 * reference's defaults do, so that pointer aliasing between V, V_tmp, CG_p
 * and evec behaves identically.
 *
 * Deliberate differences from the reference, none of which changes a result
 * on the default path:
 *  - MPI is absent (the reference's without-mpi configuration).
 *  - Printing is reduced to the per-iteration report of
 *    GCGE_PrintIterationInfo with print_eval = 1, written to a trace file.
 *  - The column version of GCGE_Orthonormalization never increments its
 *    reorthogonalization counter in the reference (gcge_orthogonal.c:202);
 *    here the loop is additionally capped at 1000 passes so that a case the
 *    reference would never leave cannot hang a test.
 *  - conv_type "NS" (||Ax - lambda Bx|| / (|lambda| ||Bx||)) is an extension
 *    with no counterpart in the reference.
 */
#include "gcge_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Fortran LAPACK/BLAS, LP64 (the same routines the reference binds in
 * src/gcge_ops.h:217-241). */
extern void dsyevx_(const char *jobz, const char *range, const char *uplo, int *n,
                    double *a, int *lda, double *vl, double *vu, int *il, int *iu,
                    double *abstol, int *m, double *w, double *z, int *ldz,
                    double *work, int *lwork, int *iwork, int *ifail, int *info);
extern void dgemm_(const char *ta, const char *tb, int *m, int *n, int *k,
                   double *alpha, double *a, int *lda, double *b, int *ldb,
                   double *beta, double *c, int *ldc);
extern void dsymm_(const char *side, const char *uplo, int *m, int *n,
                   double *alpha, double *a, int *lda, double *b, int *ldb,
                   double *beta, double *c, int *ldc);
extern double ddot_(int *n, double *x, int *incx, double *y, int *incy);
extern double dnrm2_(int *n, double *x, int *incx);
extern void dscal_(int *n, double *a, double *x, int *incx);
extern void daxpy_(int *n, double *a, double *x, int *incx, double *y, int *incy);
extern void dcopy_(int *n, double *x, int *incx, double *y, int *incy);

/* ------------------------------------------------------------------------ */
/* CSR backend (app/csr/gcge_app_csr.c)                                      */
/* ------------------------------------------------------------------------ */

/* CSR_MatDotVec, gcge_app_csr.c:297-323: y = 0, then row-sequential sums. */
void orc_csr_matvec(const orc_csr *A, const double *x, double *y)
{
    int i, j;
    memset(y, 0, (size_t)A->n * sizeof(double));
    for (i = 0; i < A->n; i++) {
        int s = A->rowptr[i], e = A->rowptr[i + 1];
        for (j = s; j < e; j++) y[i] += A->val[j] * x[A->col[j]];
    }
}

void orc_csr_matmultivec(const orc_csr *A, const double *X, int ldx,
                         double *Y, int ldy, int k)
{
    int c;
    for (c = 0; c < k; c++) orc_csr_matvec(A, X + (size_t)c * ldx, Y + (size_t)c * ldy);
}

/* CSR_VecAxpby, gcge_app_csr.c:330-365 (same branch structure). */
void orc_vec_axpby(double a, const double *x, double b, double *y, int n)
{
    int i;
    if (a == 0.0) {
        for (i = 0; i < n; i++) y[i] = b * y[i];
    } else if (b == 0.0) {
        if (a == 1.0) {
            for (i = 0; i < n; i++) y[i] = x[i];
        } else {
            for (i = 0; i < n; i++) y[i] = a * x[i];
        }
    } else {
        for (i = 0; i < n; i++) y[i] = a * x[i] + b * y[i];
    }
}

/* CSR_VecInnerProd, gcge_app_csr.c:368-378: sequential sum from 0. */
double orc_vec_dot(const double *x, const double *y, int n)
{
    int i;
    double s = 0.0;
    for (i = 0; i < n; i++) s += x[i] * y[i];
    return s;
}

/* CSR_VecSetRandomValue, gcge_app_csr.c:284-294: rand() / (RAND_MAX + 1). */
static void vec_random(double *x, int n)
{
    int j;
    for (j = 0; j < n; j++) x[j] = ((double)rand()) / ((double)RAND_MAX + 1);
}

/* CSR_SetDirichletBoundary, gcge_app_csr.c:130-183: zero row i of every
 * vector when row i of B has no entry above 1e-20 and row i of A has exactly
 * one. */
static void csr_dirichlet(double **vecs, int nev, const orc_csr *A, const orc_csr *B)
{
    int i, j, k;
    const double tol = 1e-20;
    if (A == NULL || B == NULL) return;
    for (i = 0; i < A->n; i++) {
        int b_empty = 1, a_count = 0;
        for (j = B->rowptr[i]; j < B->rowptr[i + 1]; j++)
            if (fabs(B->val[j]) > tol) { b_empty = 0; break; }
        if (!b_empty) continue;
        for (j = A->rowptr[i]; j < A->rowptr[i + 1]; j++)
            if (fabs(A->val[j]) > tol) a_count++;
        if (a_count == 1)
            for (k = 0; k < nev; k++) vecs[k][i] = 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Solver state (GCGE_PARA run-time fields + GCGE_WORKSPACE)                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_csr *A, *B;
    int n;
    orc_params P;
    /* run-time GCGE_PARA fields */
    int num_unlock, num_iter;
    double *res, max_res, min_res, sum_res, max_ind, min_ind;
    /* GCGE_WORKSPACE (gcge_workspace.h:26-66) */
    double **V, **V_tmp, **CG_p, **evec;
    int V_size, V_tmp_size, CG_p_size;
    int max_dim_x, dim_x, dim_xp, dim_xpw, last_dim_xpw, last_dim_x;
    int unconv_bs, num_soft_locked;
    double *subspace_matrix, *subspace_evec, *subspace_dtmp, *eval;
    int *subspace_itmp, *unlock;
    orc_info *info;
    FILE *trace;
} S_t;

static void matvec(S_t *S, const orc_csr *M, const double *x, double *y)
{
    orc_csr_matvec(M, x, y);
    S->info->spmv_count++;
}

/* ------------------------------------------------------------------------ */
/* Default multivector operations (src/gcge_ops.c)                           */
/* ------------------------------------------------------------------------ */

/* GCGE_Default_MatDotMultiVec, gcge_ops.c:115-129 */
static void mv_matdot(S_t *S, const orc_csr *M, double **x, double **y, const int *s, const int *e)
{
    int i, k = e[0] - s[0];
    for (i = 0; i < k; i++) matvec(S, M, x[s[0] + i], y[s[1] + i]);
}

/* GCGE_Default_MultiVecAxpby, gcge_ops.c:132-146 */
static void mv_axpby(S_t *S, double a, double **x, double b, double **y, const int *s, const int *e)
{
    int i, k = e[0] - s[0];
    for (i = 0; i < k; i++) orc_vec_axpby(a, x[s[0] + i], b, y[s[1] + i], S->n);
}

/* GCGE_Default_MultiVecLinearComb, gcge_ops.c:163-193: the first term is
 * written with b = 0, the others accumulated with b = 1. */
static void mv_lincomb(S_t *S, double **x, double **y, const int *s, const int *e,
                       const double *a, int lda)
{
    int iy, ix;
    for (iy = s[1]; iy < e[1]; iy++) {
        double *ys = y[iy];
        orc_vec_axpby(a[(iy - s[1]) * lda], x[s[0]], 0.0, ys, S->n);
        for (ix = s[0] + 1; ix < e[0]; ix++)
            orc_vec_axpby(a[(iy - s[1]) * lda + ix - s[0]], x[ix], 1.0, ys, S->n);
    }
}

/* GCGE_Default_MultiVecInnerProd, gcge_ops.c:198-262.  "sym": upper triangle
 * column by column (row count from the W range), then mirrored. */
static void mv_innerprod(S_t *S, double **V, double **W, double *a, const char *is_sym,
                         const int *s, const int *e, int lda)
{
    int iv, iw;
    double *p = a;
    if (strcmp(is_sym, "sym") == 0) {
        int len = e[1] - s[1];
        for (iw = 0; iw < len; iw++) {
            for (iv = 0; iv <= iw; iv++, p++)
                *p = orc_vec_dot(V[iv + s[0]], W[iw + s[1]], S->n);
            p += lda - (iw + 1);
        }
        for (iw = 0; iw < len; iw++)
            for (iv = iw + 1; iv < len; iv++) a[iw * lda + iv] = a[iv * lda + iw];
    } else {
        int skip = lda - (e[0] - s[0]);
        for (iw = s[1]; iw < e[1]; iw++) {
            for (iv = s[0]; iv < e[0]; iv++, p++) *p = orc_vec_dot(V[iv], W[iw], S->n);
            p += skip;
        }
    }
}

/* GCGE_Default_MultiVecSwap, gcge_ops.c:285-300: pointer exchange. */
static void mv_swap(double **V1, double **V2, const int *s, const int *e)
{
    int i, k = e[0] - s[0];
    for (i = 0; i < k; i++) {
        double *t = V1[s[0] + i];
        V1[s[0] + i] = V2[s[1] + i];
        V2[s[1] + i] = t;
    }
}

/* Public kernel-level restatements. */
void orc_multivec_innerprod(const double *V, int ldv, int p, const double *W, int ldw, int q,
                            double *a, int lda)
{
    int i, j, n = ldv < ldw ? ldv : ldw;
    (void)n;
    for (j = 0; j < q; j++)
        for (i = 0; i < p; i++) {
            /* length is the common row count; callers pass ldv == ldw == n */
            a[(size_t)j * lda + i] = orc_vec_dot(V + (size_t)i * ldv, W + (size_t)j * ldw, ldv);
        }
}

void orc_multivec_lincomb(const double *X, int ldx, int p, double *Y, int ldy, int q,
                          const double *a, int lda, int n)
{
    int iy, ix;
    for (iy = 0; iy < q; iy++) {
        double *ys = Y + (size_t)iy * ldy;
        orc_vec_axpby(a[(size_t)iy * lda], X, 0.0, ys, n);
        for (ix = 1; ix < p; ix++)
            orc_vec_axpby(a[(size_t)iy * lda + ix], X + (size_t)ix * ldx, 1.0, ys, n);
    }
}

/* ------------------------------------------------------------------------ */
/* Iteration report (GCGE_PrintIterationInfo, gcge_para.c:390-437)           */
/* ------------------------------------------------------------------------ */
static void print_iteration(S_t *S)
{
    int i;
    const char *ct;
    if (!S->trace) return;
    fprintf(S->trace, "num_unlock(%3d) = %3d; max_res(%3d) = %e; min_res(%3d) = %e;\n",
            S->num_iter, S->num_unlock, S->num_iter, S->max_res, S->num_iter, S->min_res);
    if (strcmp(S->P.conv_type, "R") == 0) ct = "relative_res";
    else if (strcmp(S->P.conv_type, "O") == 0) ct = "omega_res";
    else ct = "abosolute_res";
    for (i = 0; i < S->P.nev; i++) {
        fprintf(S->trace, "eval(%d,%3d) = %20.15e ; %s(%d,%3d) =  %20.15e ;",
                S->num_iter, i + 1, S->eval[i], ct, S->num_iter, i + 1, S->res[i]);
        if (S->res[i] < S->P.ev_tol) fprintf(S->trace, " (converged)\n");
        else if (i == S->max_ind) fprintf(S->trace, " (max)\n");
        else if (i == S->min_ind) fprintf(S->trace, " (min)\n");
        else fprintf(S->trace, "\n");
    }
    fflush(S->trace);
}

/* ------------------------------------------------------------------------ */
/* Orthonormalization (src/gcge_orthogonal.c)                                */
/* ------------------------------------------------------------------------ */

/* GCGE_Orthonormalization, gcge_orthogonal.c:112-305. */
static void orth_gs(S_t *S, double **V, int start, int *end, const orc_csr *B,
                    double **V_tmp, double *d_tmp)
{
    int current, j, reorth_count, guard;
    int ms[2], me[2];
    double vin = 0.0, vout = 0.0, norm_value = 0.0, ipv = 0.0, ratio = 1.0;
    double sum_res = S->sum_res;
    const double crit = S->P.criterion_tol;
    const int n = S->n;
    for (current = start; current < *end; ++current) {
        reorth_count = 0;
        if (sum_res > 0.0 && sum_res < crit) {
            /* column version, :139-208 */
            double *vc = V[current];
            if (B == NULL) {
                norm_value = orc_vec_dot(vc, vc, n);
            } else {
                matvec(S, B, vc, V_tmp[1]);
                norm_value = orc_vec_dot(vc, V_tmp[1], n);
            }
            vout = sqrt(norm_value);
            guard = 0;
            do {
                for (j = 0; j < current; j++) {
                    if (B == NULL) {
                        ipv = orc_vec_dot(vc, V[j], n);
                    } else {
                        matvec(S, B, vc, V_tmp[1]);
                        ipv = orc_vec_dot(V[j], V_tmp[1], n);
                    }
                    orc_vec_axpby(-ipv, V[j], 1.0, vc, n);
                }
                vin = vout;
                if (B == NULL) {
                    norm_value = orc_vec_dot(vc, vc, n);
                } else {
                    matvec(S, B, vc, V_tmp[1]);
                    norm_value = orc_vec_dot(vc, V_tmp[1], n);
                }
                vout = sqrt(norm_value);
                ratio = vout / vin;
                guard++;
            } while (ratio < S->P.reorth_tol && reorth_count < S->P.max_reorth_time &&
                     vout > S->P.orth_zero_tol && guard < 1000);
        } else {
            /* block version, :209-273 */
            do {
                if (B == NULL) orc_vec_axpby(1.0, V[current], 0.0, V_tmp[0], n);
                else matvec(S, B, V[current], V_tmp[0]);
                ms[0] = 0; me[0] = current + 1; ms[1] = 0; me[1] = 1;
                mv_innerprod(S, V, V_tmp, d_tmp, "ns", ms, me, current + 1);
                if (current != 0) {
                    ms[0] = 0; me[0] = current; ms[1] = 0; me[1] = 1;
                    mv_lincomb(S, V, V_tmp, ms, me, d_tmp, current);
                    orc_vec_axpby(-1.0, V_tmp[0], 1.0, V[current], n);
                }
                vin = sqrt(d_tmp[current]);
                {
                    int one = 1, len = current;
                    d_tmp[current] = d_tmp[current] - ddot_(&len, d_tmp, &one, d_tmp, &one);
                }
                vout = sqrt(d_tmp[current]);
                ratio = vout / vin;
                reorth_count++;
            } while (ratio < S->P.reorth_tol && reorth_count < S->P.max_reorth_time &&
                     vout > S->P.orth_zero_tol);
        }
        if (vout > S->P.orth_zero_tol) {
            orc_vec_axpby(0.0, V[current], 1 / vout, V[current], n);
        } else {
            ms[0] = current; ms[1] = *end - 1; me[0] = current + 1; me[1] = *end;
            mv_swap(V, V, ms, me);
            (*end)--;
            current--;
        }
    }
}

/* GCGE_SubOrthonormalization, gcge_orthogonal.c:1880-1929:
 * V2 = V2 - V1 (V1^T B V2). */
static void sub_orth(S_t *S, double **V, const int *start, const int *end, const orc_csr *B,
                     double **V_tmp, double *d)
{
    int len1 = end[0] - start[0], len2 = end[1] - start[1];
    int ms[2], me[2];
    ms[0] = start[1]; me[0] = end[1]; ms[1] = 0; me[1] = len2;
    if (B == NULL) mv_axpby(S, 1.0, V, 0.0, V_tmp, ms, me);
    else mv_matdot(S, B, V, V_tmp, ms, me);
    ms[0] = start[0]; me[0] = end[0]; ms[1] = 0; me[1] = len2;
    mv_innerprod(S, V, V_tmp, d, "ns", ms, me, len1);
    mv_lincomb(S, V, V_tmp, ms, me, d, len1);
    ms[0] = 0; me[0] = len2; ms[1] = start[1]; me[1] = end[1];
    mv_axpby(S, -1.0, V_tmp, 1.0, V, ms, me);
}

/* GCGE_SubOrthonormalizationSelfBGS, gcge_orthogonal.c:2035-2115. */
static void self_bgs(S_t *S, double **V, int start, int *end, const orc_csr *B)
{
    int current, j, ms[2], me[2];
    double *d = S->subspace_dtmp, norm_value;
    for (current = start; current < *end; current++) {
        if (B == NULL) orc_vec_axpby(1.0, V[current], 0.0, S->V_tmp[0], S->n);
        else matvec(S, B, V[current], S->V_tmp[0]);
        ms[0] = current; me[0] = *end; ms[1] = 0; me[1] = 1;
        mv_innerprod(S, V, S->V_tmp, d, "ns", ms, me, me[0] - ms[0]);
        norm_value = sqrt(d[0]);
        if (norm_value > S->P.orth_zero_tol) {
            orc_vec_axpby(0.0, V[current], 1.0 / norm_value, V[current], S->n);
            for (j = current + 1; j < *end; j++)
                orc_vec_axpby(-d[j - current] / norm_value, V[current], 1.0, V[j], S->n);
        } else {
            ms[0] = current; ms[1] = (*end) - 1; me[0] = current + 1; me[1] = *end;
            mv_swap(V, V, ms, me);
            (*end)--;
            current--;
        }
    }
}

/* GCGE_MultiOrthonormalization, gcge_orthogonal.c:1945-2033. */
static void multi_orth(S_t *S, double **V, int start, int *end, const orc_csr *B)
{
    int length = *end - start;
    int length_1 = length / 2, length_2 = length - length_1, mid = start + length_1;
    int i, iter = 0, ms[2], me[2];
    double value_inner = 1.0, t;
    const double tol = S->P.scbgs_reorth_tol;
    if (length <= S->P.max_direct_orth_length) {
        self_bgs(S, V, start, end, B);
        return;
    }
    multi_orth(S, V, start, &mid, B);
    if (mid + length_2 < *end) {
        int new_end_2 = mid + length_2, n_zero = *end - new_end_2;
        ms[0] = mid; me[0] = mid + n_zero; ms[1] = new_end_2; me[1] = *end;
        mv_swap(V, V, ms, me);
        *end = new_end_2;
    }
    if (mid - start > 0) {
        while (value_inner > tol) {
            iter++;
            ms[0] = start; me[0] = mid; ms[1] = mid; me[1] = *end;
            sub_orth(S, V, ms, me, B, S->V_tmp, S->subspace_dtmp);
            value_inner = 0.0;
            /* the reference sums length_1*length_2 entries with the
             * pre-deflation length_1 (:2014) */
            for (i = 0; i < length_1 * length_2; i++) {
                t = S->subspace_dtmp[i];
                value_inner += t * t;
            }
            value_inner = sqrt(value_inner);
            if (iter >= S->P.max_reorth_time) break;
        }
        if (value_inner > tol && S->trace)
            fprintf(S->trace, "in GCGE_SubOrthonormalization, V1: %d-%d, V2: %d-%d, "
                    "reorth_count: %d, value_inner: %e\n",
                    start, mid - 1, mid, *end, S->P.max_reorth_time, value_inner);
    }
    multi_orth(S, V, mid, end, B);
}

/* GCGE_StableMultiOrthonormalization, gcge_orthogonal.c:1833-1877. */
static void stable_multi_orth(S_t *S, double **V, int start, int *end, const orc_csr *B)
{
    int ms[2], me[2], iter = 0, i;
    double value_inner = 1.0, t;
    ms[0] = 0; me[0] = start; ms[1] = start; me[1] = *end;
    while (value_inner > S->P.scbgs_reorth_tol) {
        iter++;
        sub_orth(S, V, ms, me, B, S->V_tmp, S->subspace_dtmp);
        value_inner = 0.0;
        for (i = 0; i < start * (*end - start); i++) {
            t = S->subspace_dtmp[i];
            value_inner += t * t;
        }
        value_inner = sqrt(value_inner);
        if (iter >= S->P.max_reorth_time) break;
    }
    multi_orth(S, V, start, end, B);
}

/* GCGE_MultiOrthonormalizationInSubspace, gcge_orthogonal.c:2130-2204
 * (Euclidean; the second product updates ldV rows, as the reference does). */
static void multi_orth_insub(S_t *S, double *Vd, int ldV, int nrows, int start, int *end)
{
    int length = *end - start;
    int length_1 = length / 2, length_2 = length - length_1, mid = start + length_1;
    int rc, one = 1;
    double alpha, beta, v;
    if (length == 1) {
        v = dnrm2_(&nrows, Vd + (size_t)start * ldV, &one);
        if (v > S->P.orth_zero_tol) {
            double sc = 1.0 / v;
            dscal_(&nrows, &sc, Vd + (size_t)start * ldV, &one);
        } else {
            *end = start;
        }
        return;
    }
    multi_orth_insub(S, Vd, ldV, nrows, start, &mid);
    if (mid + length_2 < *end) {
        int new_end_2 = mid + length_2, n_zero = *end - new_end_2;
        length_1 -= n_zero;
        memcpy(Vd + (size_t)mid * ldV, Vd + (size_t)new_end_2 * ldV,
               (size_t)n_zero * ldV * sizeof(double));
        *end = new_end_2;
    }
    if (length_1 > 0) {
        for (rc = 0; rc < S->P.max_reorth_time; rc++) {
            alpha = 1.0; beta = 0.0;
            dgemm_("T", "N", &length_1, &length_2, &nrows, &alpha, Vd + (size_t)start * ldV, &ldV,
                   Vd + (size_t)mid * ldV, &ldV, &beta, S->subspace_dtmp, &length_1);
            alpha = -1.0; beta = 1.0;
            dgemm_("N", "N", &ldV, &length_2, &length_1, &alpha, Vd + (size_t)start * ldV, &ldV,
                   S->subspace_dtmp, &length_1, &beta, Vd + (size_t)mid * ldV, &ldV);
        }
    }
    multi_orth_insub(S, Vd, ldV, nrows, mid, end);
}

/* The dense array operations of src/gcge_ops.c:361-429 (BLAS-1). */
static double arr_norm(double *x, int n)
{
    int inc = 1;
    return dnrm2_(&n, x, &inc);
}

static double arr_dot(double *x, double *y, int n)
{
    int inc = 1;
    return ddot_(&n, x, &inc, y, &inc);
}

static void arr_axpby(double a, double *x, double b, double *y, int n)
{
    int inc = 1;
    if (b != 1.0) dscal_(&n, &b, y, &inc);
    daxpy_(&n, &a, x, &inc, y, &inc);
}

static void arr_copy(double *x, double *y, int n)
{
    int inc = 1;
    dcopy_(&n, x, &inc, y, &inc);
}

static void arr_scale(double a, double *x, int n)
{
    int inc = 1;
    dscal_(&n, &a, x, &inc);
}

/* GCGE_OrthonormalizationInSubspace, gcge_orthogonal.c:581-645
 * (p_orth_type "gs" and any name the reference does not list). */
static void gs_insub(S_t *S, double *Vd, int ldV, int nrows, int start, int *end)
{
    int current, idx, reorth_count;
    double vin, vout, ratio, ip;
    for (current = start; current < *end; ++current) {
        double *vc = Vd + (size_t)current * ldV;
        vout = arr_norm(vc, nrows);
        if (current != 0) {
            reorth_count = 0;
            do {
                vin = vout;
                for (idx = 0; idx < current; idx++) {
                    ip = arr_dot(Vd + (size_t)idx * ldV, vc, nrows);
                    arr_axpby(-ip, Vd + (size_t)idx * ldV, 1.0, vc, nrows);
                }
                vout = arr_norm(vc, nrows);
                ratio = vout / vin;
                reorth_count++;
            } while (ratio < S->P.reorth_tol && reorth_count < S->P.max_reorth_time && vout > S->P.orth_zero_tol);
        }
        if (vout > S->P.orth_zero_tol) {
            arr_scale(1.0 / vout, vc, nrows);
        } else {
            if (current != *end - 1) arr_copy(Vd + (size_t)(*end - 1) * ldV, vc, nrows);
            (*end)--;
            current--;
        }
    }
}

/* the self sweep of GCGE_BOrthonormalizationInSubspace (:1538-1575) and
 * GCGE_SCBOrthonormalizationInSubspace (:1674-1714) */
static void self_insub(S_t *S, double *Vd, int ldV, int nrows, int start, int *end)
{
    int current, j, len, one = 1;
    double alpha = 1.0, beta = 0.0, norm, tol = S->P.orth_zero_tol, *d = S->subspace_dtmp;
    for (current = start; current < *end; ++current) {
        double *vc = Vd + (size_t)current * ldV;
        len = *end - current;
        dgemm_("T", "N", &len, &one, &nrows, &alpha, vc, &ldV, vc, &ldV, &beta, d, &len);
        norm = sqrt(d[0]);
        if (norm > tol) {
            arr_scale(1.0 / norm, vc, nrows);
            for (j = current + 1; j < *end; j++)
                if (fabs(d[j - current] / norm) > 10.0 * tol)
                    arr_axpby(-d[j - current] / norm, vc, 1.0, Vd + (size_t)j * ldV, nrows);
        } else {
            if (current != *end - 1) arr_copy(Vd + (size_t)(*end - 1) * ldV, vc, nrows);
            (*end)--;
            current--;
        }
    }
}

/* GCGE_BOrthonormalizationInSubspace, gcge_orthogonal.c:1501-1578. */
static void bgs_insub(S_t *S, double *Vd, int ldV, int nrows, int start, int *end)
{
    int pass, current, j, len, one = 1;
    double alpha = 1.0, beta = 0.0, tol = S->P.orth_zero_tol, *d = S->subspace_dtmp;
    for (pass = 0; pass < 2; pass++) {
        len = *end - start;
        for (current = 0; current < start; ++current) {
            dgemm_("T", "N", &len, &one, &nrows, &alpha, Vd + (size_t)start * ldV, &ldV, Vd + (size_t)current * ldV,
                   &ldV, &beta, d, &len);
            for (j = start; j < *end; j++)
                if (fabs(d[j - start]) > 10.0 * tol)
                    arr_axpby(-d[j - start], Vd + (size_t)current * ldV, 1.0, Vd + (size_t)j * ldV, nrows);
        }
        self_insub(S, Vd, ldV, nrows, start, end);
    }
}

/* GCGE_SCBOrthonormalizationInSubspace, gcge_orthogonal.c:1619-1718. */
static void scbgs_insub(S_t *S, double *Vd, int ldV, int nrows, int start, int *end)
{
    int rc, len;
    double alpha, beta, *d = S->subspace_dtmp;
    if (start != 0) {
        for (rc = 0; rc < S->P.max_reorth_time; rc++) {
            len = *end - start;
            alpha = 1.0;
            beta = 0.0;
            dgemm_("T", "N", &start, &len, &nrows, &alpha, Vd, &ldV, Vd + (size_t)start * ldV, &ldV, &beta, d,
                   &start);
            alpha = -1.0;
            beta = 1.0;
            dgemm_("N", "N", &ldV, &len, &start, &alpha, Vd, &ldV, d, &start, &beta, Vd + (size_t)start * ldV, &ldV);
        }
    }
    self_insub(S, Vd, ldV, nrows, start, end);
}

/* GCGE_StableMultiOrthonormalizationInSubspace, gcge_orthogonal.c:2215-2247. */
static void stable_multi_orth_insub(S_t *S, double *Vd, int ldV, int nrows, int start, int *end)
{
    int rc, len;
    double alpha, beta;
    if (start != 0) {
        for (rc = 0; rc < S->P.max_reorth_time; rc++) {
            len = *end - start;
            alpha = 1.0; beta = 0.0;
            dgemm_("T", "N", &start, &len, &nrows, &alpha, Vd, &ldV, Vd + (size_t)start * ldV, &ldV,
                   &beta, S->subspace_dtmp, &start);
            alpha = -1.0; beta = 1.0;
            dgemm_("N", "N", &ldV, &len, &start, &alpha, Vd, &ldV, S->subspace_dtmp, &start,
                   &beta, Vd + (size_t)start * ldV, &ldV);
        }
    }
    multi_orth_insub(S, Vd, ldV, nrows, start, end);
}

/* ------------------------------------------------------------------------ */
/* Rayleigh-Ritz (src/gcge_rayleighritz.c)                                   */
/* ------------------------------------------------------------------------ */


/* B*V[xs:xe] into Y[ys:...] (MatDotVec per column), or the copy
 * VecAxpby(1.0, x, 0.0, y) when B is NULL. */
static void b_times(S_t *S, const orc_csr *B, double **X, int xs, int xe, double **Y, int ys)
{
    int j;
    for (j = xs; j < xe; j++) {
        if (B == NULL) orc_vec_axpby(1.0, X[j], 0.0, Y[ys + j - xs], S->n);
        else matvec(S, B, X[j], Y[ys + j - xs]);
    }
}

/* MGS of V[start:end] against itself with zero-column deflation: the W-self
 * loop of GCGE_BOrthonormalization (gcge_orthogonal.c:459-526), also
 * GCGE_SCBOrth_Self's inner loop (:1205-1277). */
static void w_self_mgs(S_t *S, double **V, int start, int *end, const orc_csr *B)
{
    int cur, j, ms[2], me[2];
    double *d = S->subspace_dtmp, norm_value;
    for (cur = start; cur < *end; cur++) {
        b_times(S, B, V, cur, cur + 1, S->CG_p, 0);
        ms[0] = cur; me[0] = *end; ms[1] = 0; me[1] = 1;
        mv_innerprod(S, V, S->CG_p, d, "ns", ms, me, *end - cur);
        norm_value = sqrt(d[0]);
        if (norm_value > S->P.orth_zero_tol) {
            orc_vec_axpby(0.0, V[cur], 1.0 / norm_value, V[cur], S->n);
            for (j = cur + 1; j < *end; j++) orc_vec_axpby(-d[j - cur] / norm_value, V[cur], 1.0, V[j], S->n);
        } else {
            ms[0] = cur; ms[1] = *end - 1; me[0] = cur + 1; me[1] = *end;
            mv_swap(V, V, ms, me);
            (*end)--;
            cur--;
        }
    }
}

/* GCGE_BOrthonormalization, gcge_orthogonal.c:352-528. */
static void orth_bgs(S_t *S, double **V, int start, int *end, const orc_csr *B)
{
    int nev = S->P.nev, dim_xp = S->dim_xp, i, cur, j, wl, ms[2], me[2];
    double *d = S->subspace_dtmp;
    b_times(S, B, V, 0, nev, S->evec, 0);
    b_times(S, B, V, nev, dim_xp, S->V_tmp, 0);
    for (i = 0; i < 2; i++) {
        wl = *end - start;
        for (cur = 0; cur < dim_xp; cur++) {
            double **BV = cur < nev ? S->evec : S->V_tmp;
            int bc = cur < nev ? cur : cur - nev;
            ms[0] = start; me[0] = *end; ms[1] = bc; me[1] = bc + 1;
            mv_innerprod(S, V, BV, d, "ns", ms, me, wl);
            for (j = start; j < *end; j++) orc_vec_axpby(-d[j - start], V[cur], 1.0, V[j], S->n);
        }
        w_self_mgs(S, V, start, end, B);
    }
}

/* V_tmp = B W, d = V[p0:p1]^T V_tmp (leading dimension lda), V_tmp = V[p0:p1] d,
 * W -= V_tmp, for W = V[start:end] (gcge_orthogonal.c:822-855); returns
 * sqrt(sum d[0:nsum]^2). */
static double project_out(S_t *S, double **V, int p0, int p1, int start, int end, int lda, int nsum,
                          const orc_csr *B)
{
    int wl = end - start, ms[2], me[2], i;
    double *d = S->subspace_dtmp, value_inner = 0.0, t;
    ms[0] = start; me[0] = end; ms[1] = 0; me[1] = wl;
    if (B == NULL) mv_axpby(S, 1.0, V, 0.0, S->V_tmp, ms, me);
    else mv_matdot(S, B, V, S->V_tmp, ms, me);
    ms[0] = p0; me[0] = p1; ms[1] = 0; me[1] = wl;
    mv_innerprod(S, V, S->V_tmp, d, "ns", ms, me, lda);
    for (i = 0; i < nsum; i++) {
        t = d[i];
        value_inner += t * t;
    }
    mv_lincomb(S, V, S->V_tmp, ms, me, d, lda);
    ms[0] = 0; me[0] = wl; ms[1] = start; me[1] = end;
    mv_axpby(S, -1.0, S->V_tmp, 1.0, V, ms, me);
    return sqrt(value_inner);
}

/* GCGE_CBOrthonormalization, gcge_orthogonal.c:755-963. */
static void orth_cbgs(S_t *S, double **V, int start, int *end, const orc_csr *B)
{
    int i, j, wl = *end - start, ms[2], me[2], il = 1, iu, m, lwork, info = 0, nzs;
    double *d = S->subspace_dtmp, vl = 0.0, vu = 0.0, abstol = 0.0;
    double *tmp_evec, *tmp_eval, *dwork;
    int *ifail, nonzero = 0;
    for (i = 0; i < 2; i++) {
        wl = *end - start;
        project_out(S, V, 0, start, start, *end, start, 0, B);
    }
    b_times(S, B, V, start, *end, S->V_tmp, 0);
    ms[0] = start; me[0] = *end; ms[1] = 0; me[1] = wl;
    mv_innerprod(S, V, S->V_tmp, d, "ns", ms, me, wl);
    for (j = 0; j < wl * wl; j++)
        if (d[j] > S->P.orth_zero_tol) {
            nonzero = 1;
            break;
        }
    if (!nonzero) {
        *end = start;
        return;
    }
    tmp_evec = S->subspace_evec;
    tmp_eval = tmp_evec + (size_t)wl * wl;
    dwork = d + (size_t)wl * wl;
    ifail = S->subspace_itmp + 5 * wl;
    iu = wl;
    m = iu - il + 1;
    lwork = 8 * wl;
    dsyevx_("V", "I", "U", &wl, d, &wl, &vl, &vu, &il, &iu, &abstol, &m, tmp_eval, tmp_evec, &wl, dwork, &lwork,
            S->subspace_itmp, ifail, &info);
    nzs = 0;
    for (j = 0; j < wl; j++) {
        if (tmp_eval[j] < S->P.orth_zero_tol) {
            nzs = j + 1;
        } else {
            double a = 1.0 / sqrt(tmp_eval[j]);
            int len = wl, inc = 1;
            dscal_(&len, &a, tmp_evec + (size_t)j * wl, &inc);
        }
    }
    wl -= nzs;
    /* the reference offsets and reads the eigenvectors with the new column
     * count as the leading dimension (:936-945) */
    ms[0] = start; me[0] = *end; ms[1] = 0; me[1] = wl;
    mv_lincomb(S, V, S->V_tmp, ms, me, tmp_evec + (size_t)nzs * wl, wl);
    *end = start + wl;
    ms[0] = start; me[0] = *end; ms[1] = 0; me[1] = wl;
    mv_swap(V, S->V_tmp, ms, me);
}

/* GCGE_SCBOrth_Minus, gcge_orthogonal.c:1116-1193. */
static void scbgs_minus(S_t *S, double **V, int start, int end, const orc_csr *B)
{
    int wl = end - start, iter = 0, dim_xp = S->dim_xp;
    double value_inner = 1.0;
    while (value_inner > S->P.scbgs_reorth_tol) {
        iter++;
        value_inner = project_out(S, V, dim_xp, start, start, end, start, (start - dim_xp) * wl, B);
        if (iter >= S->P.max_reorth_time) break;
    }
}

/* GCGE_SCBOrthonormalization, gcge_orthogonal.c:967-1113, with
 * GCGE_SCBOrth_Self (:1196-1279). */
static void orth_scbgs(S_t *S, double **V, int start, int *end, const orc_csr *B)
{
    int wl = *end - start, iter = 0, step, w_start, w_end, old_w_end, it, ms[2], me[2];
    double value_inner = 1.0;
    while (value_inner > S->P.scbgs_reorth_tol) {
        iter++;
        value_inner = project_out(S, V, 0, start, start, *end, start, start * wl, B);
        if (iter >= S->P.max_reorth_time) break;
    }
    step = (*end - start < S->P.w_orth_block_size) ? (*end - start) : S->P.w_orth_block_size;
    w_start = start;
    w_end = w_start + step;
    if (w_end > *end) w_end = *end;
    for (it = 0; it < S->P.scbgs_wself_max_reorth_time; it++) w_self_mgs(S, V, w_start, &w_end, B);
    while (w_end < *end) {
        w_start = w_end;
        w_end = w_start + step;
        if (w_end > *end) w_end = *end;
        old_w_end = w_end;
        scbgs_minus(S, V, w_start, w_end, B);
        for (it = 0; it < S->P.scbgs_wself_max_reorth_time; it++) w_self_mgs(S, V, w_start, &w_end, B);
        if (w_end < old_w_end) {
            *end -= old_w_end - w_end;
            ms[0] = w_end; ms[1] = *end; me[0] = old_w_end; me[1] = *end + old_w_end - w_end;
            mv_swap(V, V, ms, me);
        }
    }
}

/* GCGE_ComputeSubspaceMatrixVTAW, gcge_rayleighritz.c:35-81. */
static void rr_vtaw(S_t *S, double **V, const orc_csr *A, int start_W, int end_V, double *mat,
                    double **work)
{
    int w_length = end_V - start_W, xp_length = end_V - w_length;
    int ms[2], me[2];
    ms[0] = start_W; me[0] = end_V; ms[1] = 0; me[1] = w_length;
    mv_matdot(S, A, V, work, ms, me);
    ms[0] = 0; me[0] = xp_length; ms[1] = 0; me[1] = w_length;
    mv_innerprod(S, V, work, mat, "nonsym", ms, me, end_V);
    ms[0] = start_W; me[0] = end_V; ms[1] = 0; me[1] = w_length;
    mv_innerprod(S, V, work, mat + xp_length, "sym", ms, me, end_V);
}

/* GCGE_ComputeSubspaceMatrix, gcge_rayleighritz.c:113-200 (live #else branch). */
static void rr_matrix(S_t *S, const orc_csr *A, double **V)
{
    int i, dim_xp = S->dim_xp, ldm = S->dim_xpw;
    if (dim_xp != 0) {
        int lde = S->last_dim_xpw, dim_x = S->dim_x, dim_p = dim_xp - dim_x;
        double alpha = 1.0, beta = 0.0;
        double *pp = S->subspace_evec + (size_t)dim_x * lde;
        double *AApp = S->subspace_dtmp;
        double *PTAP = S->subspace_matrix + (size_t)dim_x * ldm + dim_x;
        dsymm_("L", "U", &lde, &dim_p, &alpha, S->subspace_matrix, &lde, pp, &lde, &beta, AApp, &lde);
        memset(S->subspace_matrix, 0, (size_t)ldm * ldm * sizeof(double));
        dgemm_("T", "N", &dim_p, &dim_p, &lde, &alpha, pp, &lde, S->subspace_dtmp, &lde, &beta,
               PTAP, &ldm);
        for (i = 0; i < dim_x; i++) S->subspace_matrix[i * ldm + i] = S->eval[i];
    }
    rr_vtaw(S, V, A, dim_xp, ldm, S->subspace_matrix + (size_t)dim_xp * ldm, S->evec);
}

/* GCGE_ComputeSubspaceEigenpairs, gcge_rayleighritz.c:364-590 (single
 * process, orth_type "B").  With opt_rr_eig_partly the pairs before the
 * first unconverged one (unlock[0]), less a cluster of equal eigenvalues
 * reaching back from it (multi_tol_for_lock), are soft-locked and only the
 * trailing block is decomposed (:380-394). */
static int rr_eigenpairs(S_t *S)
{
    int ldm = S->dim_xpw, info = 0, i;
    double *temp = S->subspace_dtmp, *dwork = temp + (size_t)ldm * ldm;
    double vl = 0.0, vu = 0.0, abstol = 0.0;
    int rr_eigen_start = S->unlock[0], il, iu, m, lwork = 8 * ldm, nrows, ncols;
    int *ifail = S->subspace_itmp + 5 * ldm;
    if (S->P.opt_rr_eig_partly == 0) rr_eigen_start = 0;
    il = rr_eigen_start;
    for (i = rr_eigen_start - 1; i > -1; i--)
        if (fabs(S->eval[i] - S->eval[rr_eigen_start]) / S->eval[rr_eigen_start] < S->P.multi_tol_for_lock) il -= 1;
    rr_eigen_start = il;
    il += 1;
    S->num_soft_locked = rr_eigen_start;
    iu = (S->max_dim_x < ldm) ? S->max_dim_x : ldm;
    m = iu - il + 1;
    for (i = 0; i < S->dim_xpw - S->dim_xp; i++)
        memset(S->subspace_matrix + (size_t)(S->dim_xp + i) * ldm, 0,
               (size_t)rr_eigen_start * sizeof(double));
    memcpy(temp, S->subspace_matrix, (size_t)ldm * ldm * sizeof(double));
    nrows = ldm; ncols = ldm;
    nrows -= rr_eigen_start; il -= rr_eigen_start; iu -= rr_eigen_start;
    dsyevx_("V", "I", "U", &nrows, temp + (size_t)rr_eigen_start * ncols + rr_eigen_start, &ncols,
            &vl, &vu, &il, &iu, &abstol, &m, S->eval + rr_eigen_start,
            S->subspace_evec + (size_t)rr_eigen_start * ldm + rr_eigen_start, &ldm, dwork, &lwork,
            S->subspace_itmp, ifail, &info);
    return info;
}

/* GCGE_CheckEvalMultiplicity, gcge_eigsol.c:696-718 (threshold 0.2 fixed). */
static void check_multiplicity(int start, int nev, int end, int *dim_x, const double *eval)
{
    int tmp = start, i;
    for (i = start; i < end; i++) {
        if ((fabs(fabs(eval[tmp] / eval[nev]) - 1)) < 0.2) tmp += 1;
        else break;
    }
    *dim_x = tmp;
}

/* ------------------------------------------------------------------------ */
/* X, W, P (src/gcge_xpw.c) and the CG solvers (src/gcge_cg.c)               */
/* ------------------------------------------------------------------------ */

/* GCGE_ComputeX, gcge_xpw.c:47-76. */
static void compute_x(S_t *S, double **V, double *coef, double **X, int size_X, int size_V)
{
    int s[2], e[2], sx = S->num_soft_locked;
    s[0] = sx; e[0] = size_V; s[1] = sx; e[1] = size_X;
    mv_lincomb(S, V, X, s, e, coef + (size_t)sx * size_V + sx, size_V);
}

/* GCGE_BCG, gcge_cg.c:137-302.  The active set is rebuilt with
 * unlock[idx] = idx (:289-293), so later sweeps update columns
 * 0..num_unlock-1; this is reproduced. */
static void bcg_impl(S_t *S, const orc_csr *A, double **RHS, double **V, int x_start, int x_length,
                     double **CG_P, double **CG_W, double *dtmp, int *itmp, int max_it, double rate)
{
    double *rho1 = dtmp, *rho2 = rho1 + x_length, *error = rho2 + x_length;
    double *last_error = error + x_length, *ptw = last_error + x_length;
    int *unlock = itmp, num_unlock = 0, id, idx, niter;
    double alpha, beta;
    const int n = S->n;
    for (idx = 0; idx < x_length; idx++) {
        double *x = V[x_start + idx], *r = CG_W[idx], *b = RHS[idx];
        matvec(S, A, x, r);
        orc_vec_axpby(-1.0, r, 1.0, b, n);
        rho2[idx] = orc_vec_dot(b, b, n);
        unlock[idx] = idx;
        num_unlock++;
    }
    for (idx = 0; idx < x_length; idx++) error[idx] = sqrt(rho2[idx]);
    niter = 1;
    while (num_unlock > 0 && niter < max_it) {
        if (niter == 1) {
            for (id = 0; id < num_unlock; id++) {
                idx = unlock[id];
                orc_vec_axpby(1.0, RHS[idx], 0.0, CG_P[idx], n);
            }
        } else {
            for (id = 0; id < num_unlock; id++) {
                idx = unlock[id];
                beta = rho2[idx] / rho1[idx];
                orc_vec_axpby(1.0, RHS[idx], beta, CG_P[idx], n);
            }
        }
        for (id = 0; id < num_unlock; id++) {
            idx = unlock[id];
            matvec(S, A, CG_P[idx], CG_W[idx]);
            ptw[idx] = orc_vec_dot(CG_P[idx], CG_W[idx], n);
        }
        for (id = 0; id < num_unlock; id++) {
            idx = unlock[id];
            alpha = rho2[idx] / ptw[idx];
            orc_vec_axpby(alpha, CG_P[idx], 1.0, V[x_start + idx], n);
            orc_vec_axpby(-alpha, CG_W[idx], 1.0, RHS[idx], n);
            rho1[idx] = rho2[idx];
            rho2[idx] = orc_vec_dot(RHS[idx], RHS[idx], n);
        }
        for (idx = 0; idx < x_length; idx++) last_error[idx] = sqrt(fabs(rho2[idx]));
        num_unlock = 0;
        for (idx = 0; idx < x_length; idx++) {
            if (last_error[idx] / error[idx] >= rate) {
                unlock[idx] = idx;
                num_unlock++;
            }
        }
        niter++;
    }
    if (S->info) {
        S->info->bcg_calls++;
        S->info->bcg_iters_total += niter - 1;
    }
}

/* GCGE_CG, gcge_cg.c:26-113 (used only when if_use_bcg != 1). */
static void cg_impl(S_t *S, const orc_csr *A, double *b, double *x, double **V_tmp1, double **V_tmp2)
{
    double *r = V_tmp1[1], *p = V_tmp2[0], *w = V_tmp2[1];
    int niter, max_it = S->P.cg_max_it, type = S->P.cg_type;
    double rate = S->P.cg_rate, alpha, beta, rho1 = 0.0, rho2, error, last_error, pTw, wTw;
    const int n = S->n;
    matvec(S, A, x, r);
    orc_vec_axpby(1.0, b, -1.0, r, n);
    rho2 = orc_vec_dot(r, r, n);
    error = sqrt(rho2);
    niter = 1;
    last_error = error;
    while ((last_error / error >= rate) && (niter < max_it)) {
        if (niter == 1) {
            orc_vec_axpby(1.0, r, 0.0, p, n);
        } else {
            beta = rho2 / rho1;
            orc_vec_axpby(1.0, r, beta, p, n);
        }
        matvec(S, A, p, w);
        if (type == 1) {
            pTw = orc_vec_dot(p, w, n);
            alpha = rho2 / pTw;
            orc_vec_axpby(alpha, p, 1.0, x, n);
            orc_vec_axpby(-alpha, w, 1.0, r, n);
            rho1 = rho2;
            rho2 = orc_vec_dot(r, r, n);
        } else {
            pTw = orc_vec_dot(p, w, n);
            wTw = orc_vec_dot(w, w, n);
            alpha = rho2 / pTw;
            orc_vec_axpby(alpha, p, 1.0, x, n);
            orc_vec_axpby(-alpha, w, 1.0, r, n);
            rho1 = rho2;
            rho2 = rho1 - 2.0 * alpha * pTw + alpha * alpha * wTw;
        }
        last_error = sqrt(rho2);
        niter++;
    }
}

/* GCGE_ComputeW, gcge_xpw.c:118-217. */
static void compute_w(S_t *S, const orc_csr *A, const orc_csr *B, double **V, const double *eval)
{
    int w_start = S->dim_xp, w_length = S->unconv_bs, idx;
    const int n = S->n;
    for (idx = 0; idx < w_length; idx++) {
        double *rhs = S->V_tmp[idx];
        int xi = S->unlock[idx], wi = w_start + idx;
        double *xc = V[xi], *wc = V[wi];
        if (B == NULL) orc_vec_axpby(1.0, xc, 0.0, rhs, n);
        else matvec(S, B, xc, rhs);
        if (fabs(eval[xi]) <= 1.0) {
            orc_vec_axpby(0.0, rhs, eval[xi], rhs, n);
            orc_vec_axpby(1.0, xc, 0.0, wc, n);
        } else {
            orc_vec_axpby(1.0 / eval[xi], xc, 0.0, wc, n);
        }
    }
    if (S->P.if_use_bcg == 1) {
        bcg_impl(S, A, S->V_tmp, V, w_start, w_length, S->CG_p, S->evec, S->subspace_dtmp,
                 S->subspace_itmp, S->P.cg_max_it, S->P.cg_rate);
    } else {
        for (idx = 0; idx < w_length; idx++)
            cg_impl(S, A, S->V_tmp[idx], V[w_start + idx], S->V_tmp, S->evec);
    }
}

/* GCGE_ComputeP, gcge_xpw.c:286-425 ("multi" subspace orthogonalization). */
static void compute_p(S_t *S, double *coef, double **V)
{
    int ldc = S->dim_xpw, p_start = S->dim_x, p_ncols = S->unconv_bs;
    int xx_nrows = S->last_dim_x, cp, p_end, nev = S->P.nev;
    int ms[2], me[2];
    double *px = coef + (size_t)p_start * ldc, *pp = px + xx_nrows, *xp = coef + xx_nrows;
    int orth_start, orth_length, p_end_in_V_tmp;
    memset(px, 0, (size_t)p_ncols * ldc * sizeof(double));
    for (cp = 0; cp < p_ncols; cp++) {
        int cx = S->unlock[cp];
        memcpy(pp + (size_t)cp * ldc, xp + (size_t)cx * ldc, (size_t)(ldc - xx_nrows) * sizeof(double));
    }
    orth_start = S->num_soft_locked;
    orth_length = ldc - orth_start;
    p_start -= orth_start;
    p_end = p_start + p_ncols;
    if (strcmp(S->P.p_orth_type, "scbgs") == 0)
        scbgs_insub(S, coef + (size_t)orth_start * ldc + orth_start, ldc, orth_length, p_start, &p_end);
    else if (strcmp(S->P.p_orth_type, "bgs") == 0)
        bgs_insub(S, coef + (size_t)orth_start * ldc + orth_start, ldc, orth_length, p_start, &p_end);
    else if (strcmp(S->P.p_orth_type, "multi") == 0)
        stable_multi_orth_insub(S, coef + (size_t)orth_start * ldc + orth_start, ldc, orth_length,
                                p_start, &p_end);
    else
        gs_insub(S, coef + (size_t)orth_start * ldc + orth_start, ldc, orth_length, p_start, &p_end);
    p_ncols = p_end - p_start;
    p_end += orth_start;
    p_end_in_V_tmp = p_end - nev;
    ms[0] = orth_start; me[0] = ldc; ms[1] = 0; me[1] = p_end_in_V_tmp;
    mv_lincomb(S, V, S->V_tmp, ms, me, coef + (size_t)nev * ldc + orth_start, ldc);
    ms[0] = nev; ms[1] = 0; me[0] = p_end; me[1] = p_end_in_V_tmp;
    mv_swap(V, S->V_tmp, ms, me);
    S->dim_xp = p_end;
    (void)p_ncols;
}

/* GCGE_CheckConvergence, gcge_eigsol.c:559-686. */
static void check_convergence(S_t *S, const orc_csr *A, const orc_csr *B, const double *eval,
                              double **evec)
{
    int i, j, k, *unlock = S->unlock, unconv_bs = 0, max_conv_idx = 0;
    int max_ind = 0, min_ind = 0, num_unlock = S->num_unlock;
    double res_norm, evec_norm, residual, max_res = 0.0, min_res = 0.0, sum_res = 0.0;
    double *tmp1 = S->V_tmp[0], *tmp2 = S->V_tmp[1];
    const int n = S->n;
    const char *ct = S->P.conv_type;
    for (j = 0; j < num_unlock; j++) {
        double *ev;
        i = unlock[j];
        ev = evec[i];
        matvec(S, A, ev, tmp1);
        if (B == NULL) {
            orc_vec_axpby(-eval[i], ev, 1.0, tmp1, n);
        } else {
            matvec(S, B, ev, tmp2);
            orc_vec_axpby(-eval[i], tmp2, 1.0, tmp1, n);
        }
        res_norm = sqrt(orc_vec_dot(tmp1, tmp1, n));
        evec_norm = sqrt(orc_vec_dot(ev, ev, n));
        if (strcmp(ct, "R") == 0) {
            residual = res_norm / evec_norm / (1.0 > fabs(eval[i]) ? 1.0 : fabs(eval[i]));
        } else if (strcmp(ct, "O") == 0) {
            residual = res_norm / evec_norm / (S->P.conv_omega_norm + fabs(eval[i]));
        } else if (strcmp(ct, "NS") == 0) {
            double bx_norm = (B == NULL) ? evec_norm : sqrt(orc_vec_dot(tmp2, tmp2, n));
            residual = res_norm / (fabs(eval[i]) * bx_norm);
        } else {
            residual = res_norm / evec_norm;
        }
        S->res[i] = residual;
        sum_res += residual;
        if (j == 0) {
            max_res = residual;
            min_res = residual;
        } else {
            if (residual > max_res) { max_res = residual; max_ind = i; }
            if (residual < min_res) { min_res = residual; min_ind = i; }
        }
        if (residual > S->P.ev_tol) {
            unlock[unconv_bs] = i;
            unconv_bs += 1;
            if (j > max_conv_idx + 8) {
                for (k = j + 1; k < num_unlock; k++) {
                    unlock[unconv_bs] = unlock[k];
                    unconv_bs += 1;
                }
                j = num_unlock + 1;
            }
        } else {
            max_conv_idx = j;
        }
    }
    S->num_unlock = unconv_bs;
    S->unconv_bs = (unconv_bs < S->P.block_size) ? unconv_bs : S->P.block_size;
    S->max_res = max_res;
    S->max_ind = max_ind;
    S->min_res = min_res;
    S->min_ind = min_ind;
    S->sum_res = sum_res;
    S->num_iter += 1;
}

/* ------------------------------------------------------------------------ */
/* Parameters and workspace                                                  */
/* ------------------------------------------------------------------------ */

/* GCGE_PARA_Create, gcge_para.c:25-99 (default-path fields). */
void orc_default_params(orc_params *p)
{
    memset(p, 0, sizeof(*p));
    p->nev = 1;
    p->ev_max_it = 30;
    p->block_size = 0;
    p->given_init_evec = 0;
    p->ev_tol = 1e-8;
    p->conv_type = "R";
    p->orth_type = "B";
    p->w_orth_type = "multi";
    p->p_orth_type = "multi";
    p->conv_omega_norm = 0.0;
    p->orth_zero_tol = 1e-16;
    p->reorth_tol = 0.75;
    p->scbgs_reorth_tol = 1e-16;
    p->criterion_tol = 1e-6;
    p->max_reorth_time = 3;
    p->max_direct_orth_length = 1;
    p->multi_tol_for_lock = 1e-6;
    p->cg_max_it = 20;
    p->cg_rate = 1e-2;
    p->cg_type = 1;
    p->if_use_bcg = 1;
    p->dirichlet_boundary = 0;
    p->opt_rr_eig_partly = 0;
    p->seed = 1;
    p->w_orth_block_size = 0;
    p->scbgs_wself_max_reorth_time = 1;
}

/* ------------------------------------------------------------------------ */
/* Driver: GCGE_Solve, gcge_eigsol.c:65-542                                  */
/* ------------------------------------------------------------------------ */
int orc_solve(const orc_csr *A, const orc_csr *B, const orc_params *p, const double *X0,
              double *eval_out, double *evec_out, double *res_out, orc_info *info_out,
              FILE *trace)
{
    S_t S_, *S = &S_;
    orc_info info_local;
    int nev, n, i, bs, max_dim_x, max_dim_x_tmp, max_dim_xpw, total_vecs;
    double *pool, **all_cols;
    int ms[2], me[2];
    void *Orth_mat_v, *RR_mat_v;
    const orc_csr *Orth_mat, *RR_mat;

    memset(S, 0, sizeof(*S));
    memset(&info_local, 0, sizeof(info_local));
    S->info = &info_local;
    S->trace = trace;
    S->A = A;
    S->B = B;
    S->n = n = A->n;
    S->P = *p;
    nev = S->P.nev;
    if (nev <= 0 || nev > n) return -1;
    if (strcmp(S->P.orth_type, "B") != 0) return -2;          /* "A" not restated */

    /* GCGE_PARA_Setup, gcge_para.c:357-386 */
    S->num_unlock = nev;
    if (S->P.block_size == 0) S->P.block_size = nev;
    if (S->P.w_orth_block_size == 0) S->P.w_orth_block_size = (S->P.block_size / 5 > 1) ? S->P.block_size / 5 : 1;
    S->P.orth_zero_tol += DBL_EPSILON;
    S->P.scbgs_reorth_tol += DBL_EPSILON;
    S->res = (double *)calloc((size_t)nev, sizeof(double));
    for (i = 0; i < nev; i++) S->res[i] = 1.0;
    S->num_iter = 0;
    S->sum_res = 0.0;

    /* GCGE_WORKSPACE_Setup, gcge_workspace.c:61-119 */
    bs = S->P.block_size;
    max_dim_x_tmp = (nev * 1.25 < nev + 8) ? (int)(nev * 1.25) : (nev + 8);
    max_dim_x = (max_dim_x_tmp > nev + 3) ? max_dim_x_tmp : (nev + 3);
    S->V_size = max_dim_x + 2 * bs;
    S->V_tmp_size = max_dim_x - nev + bs;
    S->CG_p_size = bs;
    max_dim_xpw = max_dim_x + 2 * bs;
    total_vecs = S->V_size + S->V_tmp_size + S->CG_p_size + nev;
    pool = (double *)calloc((size_t)total_vecs * n, sizeof(double));
    all_cols = (double **)malloc((size_t)total_vecs * sizeof(double *));
    if (!pool || !all_cols) return -5;
    for (i = 0; i < total_vecs; i++) all_cols[i] = pool + (size_t)i * n;
    S->V = all_cols;
    S->V_tmp = all_cols + S->V_size;
    S->CG_p = S->V_tmp + S->V_tmp_size;
    S->evec = S->CG_p + S->CG_p_size;
    S->max_dim_x = max_dim_x;
    S->dim_x = nev;
    S->dim_xp = 0;
    S->dim_xpw = nev;
    S->num_soft_locked = 0;
    S->unconv_bs = (nev < bs) ? nev : bs;
    S->eval = (double *)calloc((size_t)max_dim_xpw, sizeof(double));
    S->subspace_dtmp = (double *)calloc((size_t)max_dim_xpw * max_dim_xpw + 40 * (size_t)max_dim_xpw,
                                        sizeof(double));
    S->subspace_itmp = (int *)calloc(100 * (size_t)max_dim_xpw, sizeof(int));
    S->subspace_evec = (double *)calloc((size_t)max_dim_xpw * max_dim_xpw, sizeof(double));
    S->subspace_matrix = (double *)calloc((size_t)max_dim_xpw * max_dim_xpw, sizeof(double));
    S->unlock = (int *)calloc((size_t)max_dim_x, sizeof(int));
    for (i = 0; i < max_dim_x; i++) S->unlock[i] = i;

    /* orth_type "B": Orth_mat = B, RayleighRitz_mat = A (:103-114) */
    Orth_mat_v = (void *)B;
    RR_mat_v = (void *)A;
    Orth_mat = (const orc_csr *)Orth_mat_v;
    RR_mat = (const orc_csr *)RR_mat_v;

    /* initial vectors (:126-207) */
    if (S->P.given_init_evec == 0) {
        int num_init_evec = 0;
        srand(S->P.seed);
        do {
            for (i = 0; i < nev - num_init_evec; i++) vec_random(S->evec[num_init_evec + i], n);
            if (strcmp(S->P.conv_type, "O") == 0 && num_init_evec == 0) {
                /* omega-norm estimate (:135-173) */
                double omega_F = 0.0, AO_F = 0.0, t;
                for (i = 0; i < nev; i++) {
                    t = orc_vec_dot(S->evec[i], S->evec[i], n);
                    orc_vec_axpby(0.0, S->evec[i], 1.0 / sqrt(t), S->evec[i], n);
                }
                for (i = 0; i < nev; i++) omega_F += orc_vec_dot(S->evec[i], S->evec[i], n);
                ms[0] = 0; me[0] = nev; ms[1] = 0; me[1] = nev;
                mv_matdot(S, A, S->evec, S->V, ms, me);
                for (i = 0; i < nev; i++) AO_F += orc_vec_dot(S->V[i], S->V[i], n);
                S->P.conv_omega_norm = sqrt(AO_F / omega_F);
            }
            if (S->P.dirichlet_boundary == 1) csr_dirichlet(S->evec, nev, A, B);
            S->dim_x = nev;
            orth_gs(S, S->evec, num_init_evec, &S->dim_x, Orth_mat, S->V_tmp, S->subspace_dtmp);
            num_init_evec = S->dim_x;
        } while (num_init_evec < nev);
    } else {
        for (i = 0; i < nev; i++) memcpy(S->evec[i], X0 + (size_t)i * n, (size_t)n * sizeof(double));
        orth_gs(S, S->evec, 0, &S->dim_x, Orth_mat, S->V_tmp, S->subspace_dtmp);
    }

    ms[0] = 0; me[0] = nev; ms[1] = 0; me[1] = nev;
    mv_axpby(S, 1.0, S->evec, 0.0, S->V, ms, me);
    S->dim_xpw = S->dim_x;
    S->last_dim_x = S->dim_x;
    S->dim_xp = 0;
    rr_matrix(S, RR_mat, S->V);
    rr_eigenpairs(S);
    compute_x(S, S->V, S->subspace_evec, S->evec, nev, S->dim_xpw);
    check_convergence(S, A, B, S->eval, S->evec);
    print_iteration(S);

    ms[0] = S->num_soft_locked; ms[1] = S->num_soft_locked; me[0] = nev; me[1] = nev;
    mv_swap(S->V, S->evec, ms, me);
    S->dim_xp = S->dim_x;
    compute_w(S, A, B, S->V, S->eval);
    S->last_dim_xpw = S->dim_xpw;
    S->dim_xpw = S->dim_xp + S->unconv_bs;
    orth_gs(S, S->V, S->dim_x, &S->dim_xpw, Orth_mat, S->V_tmp, S->subspace_dtmp);
    S->dim_xp = nev;
    rr_matrix(S, RR_mat, S->V);
    rr_eigenpairs(S);
    compute_x(S, S->V, S->subspace_evec, S->evec, nev, S->dim_xpw);
    check_convergence(S, A, B, S->eval, S->evec);
    print_iteration(S);

    while (S->num_unlock > 0 && S->num_iter < S->P.ev_max_it) {
        compute_p(S, S->subspace_evec, S->V);
        ms[0] = S->num_soft_locked; ms[1] = S->num_soft_locked; me[0] = nev; me[1] = nev;
        mv_swap(S->V, S->evec, ms, me);
        compute_w(S, A, B, S->V, S->eval);
        S->last_dim_xpw = S->dim_xpw;
        S->dim_xpw = S->dim_xp + S->unconv_bs;
        if (strcmp(S->P.w_orth_type, "bgs") == 0)
            orth_bgs(S, S->V, S->dim_xp, &S->dim_xpw, Orth_mat);
        else if (strcmp(S->P.w_orth_type, "cbgs") == 0)
            orth_cbgs(S, S->V, S->dim_xp, &S->dim_xpw, Orth_mat);
        else if (strcmp(S->P.w_orth_type, "scbgs") == 0)
            orth_scbgs(S, S->V, S->dim_xp, &S->dim_xpw, Orth_mat);
        else if (strcmp(S->P.w_orth_type, "multi") == 0)
            stable_multi_orth(S, S->V, S->dim_xp, &S->dim_xpw, Orth_mat);
        else
            orth_gs(S, S->V, S->dim_xp, &S->dim_xpw, Orth_mat, S->V_tmp, S->subspace_dtmp);
        rr_matrix(S, RR_mat, S->V);
        rr_eigenpairs(S);
        S->last_dim_x = S->dim_x;
        check_multiplicity(nev, nev, S->max_dim_x, &S->dim_x, S->eval);
        compute_x(S, S->V, S->subspace_evec, S->evec, nev, S->dim_xpw);
        check_convergence(S, A, B, S->eval, S->evec);
        print_iteration(S);
    }

    /* soft-locked vectors live in V (gcge_eigsol.c:533-537) */
    ms[0] = 0; me[0] = S->num_soft_locked; ms[1] = 0; me[1] = S->num_soft_locked;
    if (me[0] > 0) mv_axpby(S, 1.0, S->V, 0.0, S->evec, ms, me);
    memcpy(eval_out, S->eval, (size_t)nev * sizeof(double));
    if (evec_out)
        for (i = 0; i < nev; i++)
            memcpy(evec_out + (size_t)i * n, S->evec[i], (size_t)n * sizeof(double));
    if (res_out) memcpy(res_out, S->res, (size_t)nev * sizeof(double));
    info_local.num_iter = S->num_iter;
    info_local.num_unlock = S->num_unlock;
    info_local.sum_res = S->sum_res;
    if (info_out) *info_out = info_local;

    free(S->res);
    free(S->eval);
    free(S->subspace_dtmp);
    free(S->subspace_itmp);
    free(S->subspace_evec);
    free(S->subspace_matrix);
    free(S->unlock);
    free(all_cols);
    free(pool);
    return 0;
}

int orc_solve_file(const orc_csr *A, const orc_csr *B, const orc_params *p, const double *X0,
                   double *eval, double *evec, double *res, orc_info *info, const char *trace_path)
{
    FILE *f = NULL;
    int rc;
    if (trace_path && trace_path[0]) {
        f = fopen(trace_path, "w");
        if (!f) return -10;
    }
    rc = orc_solve(A, B, p, X0, eval, evec, res, info, f);
    if (f) fclose(f);
    return rc;
}

void orc_bcg(const orc_csr *A, double *RHS, double *W, int n, int k, int cg_max_it, double cg_rate,
             int *iters_out)
{
    S_t S_, *S = &S_;
    orc_info inf;
    double **R = (double **)malloc((size_t)k * sizeof(double *));
    double **X = (double **)malloc((size_t)k * sizeof(double *));
    double **P = (double **)malloc((size_t)k * sizeof(double *));
    double **Wk = (double **)malloc((size_t)k * sizeof(double *));
    double *pool = (double *)calloc((size_t)2 * k * n, sizeof(double));
    double *dtmp = (double *)calloc((size_t)6 * k + 8, sizeof(double));
    int *itmp = (int *)calloc((size_t)k + 1, sizeof(int));
    int i;
    memset(S, 0, sizeof(*S));
    memset(&inf, 0, sizeof(inf));
    S->info = &inf;
    S->n = n;
    for (i = 0; i < k; i++) {
        R[i] = RHS + (size_t)i * n;
        X[i] = W + (size_t)i * n;
        P[i] = pool + (size_t)i * n;
        Wk[i] = pool + (size_t)(k + i) * n;
    }
    bcg_impl(S, A, R, X, 0, k, P, Wk, dtmp, itmp, cg_max_it, cg_rate);
    if (iters_out) *iters_out = inf.bcg_iters_total;
    free(R); free(X); free(P); free(Wk); free(pool); free(dtmp); free(itmp);
}
