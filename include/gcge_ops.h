/*
 * gcge_ops.h -- the GCGE operation table (the backend plug-in boundary).
 *
 * The first 35 members have the names, types and order of the reference's
 * GCGE_OPS (src/gcge_ops.h:29-215), so a backend or application written
 * against the reference reads and writes the same slots.  Conventions kept
 * from the reference:
 *   - column ranges are half-open pairs start[2]/end[2], index [0] for the
 *     first multivector argument and [1] for the second (gcge_ops.h:67-80);
 *   - dense coefficient matrices are column-major with an explicit lda;
 *   - MultiVecLinearComb: y(:, s1+j) = sum_i x(:, s0+i) a[j*lda + i]
 *     (gcge_ops.c:163-193); MultiVecInnerProd: a[j*lda + i] =
 *     <V(:, s0+i), W(:, s1+j)>, "sym" fills the upper triangle and mirrors it
 *     (gcge_ops.c:198-262);
 *   - a NULL matrix B means the identity.
 *
 * After those slots come the block-kernel slots of this build (no
 * counterpart in the reference).  The device backend (gcge_app_hip.h) fills
 * every slot; GCGE_OPS_Setup refuses a table whose block-kernel slots are
 * missing, because the solver core of this build only runs on device
 * multivectors.
 */
#ifndef GCGE_OPS_H_
#define GCGE_OPS_H_

#include "gcge_config.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GCGE_OPS_ {
    /* per-vector slots (reference gcge_ops.h:31-38) */
    void (*VecSetRandomValue)(void *vec);
    void (*MatDotVec)(void *Matrix, void *x, void *r);
    void (*VecAxpby)(GCGE_DOUBLE a, void *x, GCGE_DOUBLE b, void *y); /* y = a x + b y */
    void (*VecInnerProd)(void *x, void *y, GCGE_DOUBLE *value_ip);
    void (*VecLocalInnerProd)(void *x, void *y, GCGE_DOUBLE *value_ip);
    void (*VecCreateByVec)(void **des_vec, void *src_vec);
    void (*VecCreateByMat)(void **vec, void *mat);
    void (*VecDestroy)(void **vec);

    /* optional user linear solver (reference gcge_ops.h:43-44) */
    void (*LinearSolver)(void *Matrix, void *b, void *x, struct GCGE_OPS_ *ops);
    void *linear_solver_workspace;

    /* present in the reference, never called by its core (gcge_ops.h:48-52) */
    void (*Orthonormalization)(void **V, GCGE_INT start, GCGE_INT *end, void *B,
                               GCGE_DOUBLE orth_zero_tol, void **work);
    void (*DenseMatCreate)(void **densemat, GCGE_INT nrows, GCGE_INT ncols);
    void (*DenseMatDestroy)(void **mat);

    /* multivector slots (reference gcge_ops.h:57-86) */
    void (*MultiVecCreateByVec)(void ***multi_vec, GCGE_INT n_vec, void *vec, struct GCGE_OPS_ *ops);
    void (*MultiVecCreateByMat)(void ***multi_vec, GCGE_INT n_vec, void *mat, struct GCGE_OPS_ *ops);
    void (*MultiVecCreateByMultiVec)(void ***multi_vec, GCGE_INT n_vec, void **init_vec,
                                     struct GCGE_OPS_ *ops);
    void (*MultiVecDestroy)(void ***MultiVec, GCGE_INT n_vec, struct GCGE_OPS_ *ops);
    void (*MultiVecSetRandomValue)(void **multi_vec, GCGE_INT start, GCGE_INT n_vec,
                                   struct GCGE_OPS_ *ops);
    void (*MatDotMultiVec)(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end,
                           struct GCGE_OPS_ *ops);
    void (*MultiVecAxpby)(GCGE_DOUBLE a, void **x, GCGE_DOUBLE b, void **y, GCGE_INT *start,
                          GCGE_INT *end, struct GCGE_OPS_ *ops);
    void (*MultiVecAxpbyColumn)(GCGE_DOUBLE a, void **x, GCGE_INT col_x, GCGE_DOUBLE b, void **y,
                                GCGE_INT col_y, struct GCGE_OPS_ *ops);
    void (*MultiVecLinearComb)(void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_DOUBLE *a,
                               GCGE_INT lda, void *dmat, GCGE_INT lddmat, struct GCGE_OPS_ *ops);
    void (*MultiVecInnerProd)(void **V, void **W, GCGE_DOUBLE *a, char *is_sym, GCGE_INT *start,
                              GCGE_INT *end, GCGE_INT lda, struct GCGE_OPS_ *ops);
    void (*MultiVecSwap)(void **V_1, void **V_2, GCGE_INT *start, GCGE_INT *end,
                         struct GCGE_OPS_ *ops);
    void (*MultiVecPrint)(void **x, GCGE_INT n);
    void (*GetVecFromMultiVec)(void **V, GCGE_INT j, void **x);
    void (*RestoreVecForMultiVec)(void **V, GCGE_INT j, void **x);
    void (*SetDirichletBoundary)(void **Vecs, GCGE_INT nev, void *A, void *B);

    /* dense slots, LAPACK/BLAS shaped (reference gcge_ops.h:191-212) */
    void (*DenseMatEigenSolver)(char *jobz, char *range, char *uplo, GCGE_INT *nrows,
                                GCGE_DOUBLE *a, GCGE_INT *lda, GCGE_DOUBLE *vl, GCGE_DOUBLE *vu,
                                GCGE_INT *il, GCGE_INT *iu, GCGE_DOUBLE *abstol, GCGE_INT *nev,
                                GCGE_DOUBLE *eval, GCGE_DOUBLE *evec, GCGE_INT *lde,
                                GCGE_INT *isuppz, GCGE_DOUBLE *work, GCGE_INT *lwork,
                                GCGE_INT *iwork, GCGE_INT *liwork, GCGE_INT *ifail, GCGE_INT *info);
    void (*DenseMatDotDenseMat)(char *transa, char *transb, GCGE_INT *nrows, GCGE_INT *ncols,
                                GCGE_INT *mid, GCGE_DOUBLE *alpha, GCGE_DOUBLE *a, GCGE_INT *lda,
                                GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta, GCGE_DOUBLE *c,
                                GCGE_INT *ldc);
    void (*DenseSymMatDotDenseMat)(char *side, char *uplo, GCGE_INT *nrows, GCGE_INT *ncols,
                                   GCGE_DOUBLE *alpha, GCGE_DOUBLE *a, GCGE_INT *lda,
                                   GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta,
                                   GCGE_DOUBLE *c, GCGE_INT *ldc);
    GCGE_DOUBLE (*ArrayDotArray)(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT length);
    GCGE_DOUBLE (*ArrayNorm)(GCGE_DOUBLE *x, GCGE_INT length);
    void (*ArrayAXPBY)(GCGE_DOUBLE a, GCGE_DOUBLE *x, GCGE_DOUBLE b, GCGE_DOUBLE *y, GCGE_INT length);
    void (*ArrayCopy)(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT length);
    void (*ArrayScale)(GCGE_DOUBLE alpha, GCGE_DOUBLE *a, GCGE_INT length);

    /* ---- block-kernel slots of this build (no counterpart in the reference) ---- */

    /* y(:, s1+j) = a[j] x(:, s0+j) + b[j] y(:, s1+j), with the branch rules of
     * y = a x + b y per column (a == 0: y = b y; b == 0 and a == 1: copy;
     * b == 0: y = a x). */
    void (*MultiVecAxpbyDiag)(const GCGE_DOUBLE *a, void **x, const GCGE_DOUBLE *b, void **y,
                              GCGE_INT *start, GCGE_INT *end, struct GCGE_OPS_ *ops);
    /* out[j] = <x(:, s0+j), y(:, s1+j)> */
    void (*MultiVecInnerProdDiag)(void **x, void **y, GCGE_DOUBLE *out, GCGE_INT *start,
                                  GCGE_INT *end, struct GCGE_OPS_ *ops);
    /* y(:, s1+j) = beta y(:, s1+j) + alpha sum_i x(:, s0+i) a[j*lda + i] */
    void (*MultiVecLinearCombAxpby)(void **x, void **y, GCGE_INT *start, GCGE_INT *end,
                                    GCGE_DOUBLE *a, GCGE_INT lda, GCGE_DOUBLE alpha,
                                    GCGE_DOUBLE beta, struct GCGE_OPS_ *ops);
    /* y(:, s1+j) = mat x(:, s0+j) and out[j] = <x(:, s0+j), y(:, s1+j)> */
    void (*MatDotMultiVecInnerProdDiag)(void *mat, void **x, void **y, GCGE_INT *start,
                                        GCGE_INT *end, GCGE_DOUBLE *out, struct GCGE_OPS_ *ops);
    /* block CG step for k columns: x(:,sx+j) += alpha[j] p(:,j);
     * r(:,j) -= alpha[j] w(:,j); rr[j] = <r(:,j), r(:,j)> */
    void (*BlockCGUpdate)(void **x, GCGE_INT sx, void **p, void **w, void **r, GCGE_INT k,
                          const GCGE_DOUBLE *alpha, GCGE_DOUBLE *rr, struct GCGE_OPS_ *ops);
    /* residual norms of the pairs (lambda[j], X(:, idx[j])), j < k, computed
     * in FP64: rnorm[j] = ||A x - lambda B x||, xnorm[j] = ||x||,
     * bxnorm[j] = ||B x|| (bxnorm may be NULL) */
    void (*MultiVecResidualNorms)(void *A, void *B, void **X, const GCGE_INT *idx, GCGE_INT k,
                                  const GCGE_DOUBLE *lambda, GCGE_DOUBLE *rnorm, GCGE_DOUBLE *xnorm,
                                  GCGE_DOUBLE *bxnorm, struct GCGE_OPS_ *ops);
    /* y(:, ystart+j) = x(:, idx[j]), j < k */
    void (*MultiVecGather)(void **x, const GCGE_INT *idx, GCGE_INT k, void **y, GCGE_INT ystart,
                           struct GCGE_OPS_ *ops);
    /* copy columns x(:, s0+j) -> y(:, s1+j), converting precision if the two
     * multivectors differ in it */
    void (*MultiVecCopy)(void **x, void **y, GCGE_INT *start, GCGE_INT *end, struct GCGE_OPS_ *ops);
    /* number of rows of a multivector (for host-side checks) */
    GCGE_INT (*MultiVecNumRows)(void **x);
    /* create n_vec columns shaped by mat with the given storage precision
     * (0: FP64, 1: FP32); used for the FP32 inner solves (bcg_precision) */
    void (*MultiVecCreateByMatPrecision)(void ***multi_vec, GCGE_INT n_vec, void *mat, GCGE_INT precision,
                                         struct GCGE_OPS_ *ops);
    /* the block CG step split for fusion (same roundings as BlockCGUpdate):
     * r(:,j) -= alpha[j] w(:,j); rr[j] = <r(:,j), r(:,j)>, j < k */
    void (*BlockCGResidualUpdate)(void **w, void **r, GCGE_INT k, const GCGE_DOUBLE *alpha, GCGE_DOUBLE *rr,
                                  struct GCGE_OPS_ *ops);
    /* x(:,sx+j) += alpha[j] p(:,j) for j < kx (the deferred solution update
     * of the previous step), then p(:,j) = r(:,j) + beta[j] p(:,j) for
     * j < kp, per element in that order (kx and kp in any relation) */
    void (*BlockCGSearchUpdate)(void **x, GCGE_INT sx, void **p, void **r, GCGE_INT kx, GCGE_INT kp,
                                const GCGE_DOUBLE *alpha, const GCGE_DOUBLE *beta, struct GCGE_OPS_ *ops);
    /* the two steps of the Jacobi-preconditioned block CG (para->bcg_precond;
     * the reference's BCG has no preconditioner), with z = D^-1 r formed where
     * it is used (D the diagonal of Matrix) instead of stored.  Optional:
     * needed only when bcg_precond is set.
     * Residual step: r(:,j) -= alpha[j] w(:,j) as in BlockCGResidualUpdate,
     * rz[j] = <r(:,j), D^-1 r(:,j)>, rr[j] = <r(:,j), r(:,j)>, j < k. */
    void (*BlockPCGResidualUpdate)(void *Matrix, void **w, void **r, GCGE_INT k, const GCGE_DOUBLE *alpha,
                                   GCGE_DOUBLE *rz, GCGE_DOUBLE *rr, struct GCGE_OPS_ *ops);
    /* Search step: as BlockCGSearchUpdate with p(:,j) = D^-1 r(:,j) + beta[j] p(:,j) */
    void (*BlockPCGSearchUpdate)(void *Matrix, void **x, GCGE_INT sx, void **p, void **r, GCGE_INT kx,
                                 GCGE_INT kp, const GCGE_DOUBLE *alpha, const GCGE_DOUBLE *beta,
                                 struct GCGE_OPS_ *ops);
} GCGE_OPS;

void GCGE_OPS_Create(GCGE_OPS **ops);
GCGE_INT GCGE_OPS_Setup(GCGE_OPS *ops);
void GCGE_OPS_Free(GCGE_OPS **ops);
void GCGE_OPS_SetLinearSolverWorkspace(GCGE_OPS *ops, void *linear_solver_workspace);

#ifdef __cplusplus
}
#endif
#endif
