/*
 * gcge_app_hip.h -- the MI355X (gfx950) backend of GCGE, "app/hip".
 *
 * It plays the role app/csr plays in the reference (app/csr/gcge_app_csr.h):
 * a matrix type, vector/multivector types and the functions that fill the
 * GCGE_OPS table.  Data live in HBM: a HIP_MAT is a CSR matrix (int32 row
 * pointers and column indices, FP64 values and, on demand, an FP32 copy) and
 * a multivector is one column-major allocation with a padded leading
 * dimension.  The void ** handle of a multivector is opaque; the
 * solver reaches its columns only through the GCGE_OPS slots.
 *
 * Every function returns (or latches) errors instead of aborting: the
 * GCGE_OPS slots return void, so the first HIP or argument error is kept and
 * can be read with GCGE_HIP_GetLastError.  Once an error is latched, the
 * backend skips further device launches until GCGE_HIP_ClearError.
 */
#ifndef GCGE_APP_HIP_H_
#define GCGE_APP_HIP_H_

#include "gcge_config.h"
#include "gcge_ops.h"
#include "gcge_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct HIP_MAT_ HIP_MAT;

enum { GCGE_HIP_FP64 = 0, GCGE_HIP_FP32 = 1 };

/* ---- matrices (counterpart of CSR_MAT, gcge_app_csr.h:27-40) ---- */

/* Upload a square CSR matrix: 0-based int32 rowptr[nrows+1], col[nnz],
 * FP64 val[nnz].  Columns need not be sorted.  Returns NULL on error. */
HIP_MAT *HIP_MatCreateFromCSR(GCGE_INT nrows, GCGE_INT ncols, GCGE_INT nnz, const GCGE_INT *rowptr,
                              const GCGE_INT *col, const GCGE_DOUBLE *val);
/* Read a matrix file and upload it (HIP_MatCreateFromCSR).  Two formats:
 * the text format of CSR_ReadMatFile (app/csr/gcge_app_csr.c:185-236):
 * nrow, ncol, nnz, then rowptr, col, val, whitespace separated; and its
 * binary fast path, recognized by the magic "GCGECSR1": int32 nrow, ncol,
 * nnz, 0, then rowptr[nrow+1], col[nnz] (int32) and val[nnz] (FP64), little
 * endian, as HIP_WriteMatFile writes it. */
HIP_MAT *HIP_ReadMatFile(const char *filename);
/* A locality order of the rows of a square CSR pattern (host only, no
 * coordinates needed): perm[new position] = row.  A box grid numbered
 * lexicographically is recognized from the pattern (at least 97 % of the
 * entries within 2 grid steps in every direction) and listed in tiles of
 * 8 x 8 x 8 (3-D) or 16 x 16 (2-D) rows; any other pattern is ordered by
 * recursive halving of its graph in breadth-first order, parts of at most
 * leaf rows (0: 512) in breadth-first order.  Either way rows come in
 * compact pieces, so a 256-row block reads its own rows plus about as many
 * again (its surface) instead of whole neighbouring planes.  Returns 0, or
 * non-zero for a bad pattern. */
GCGE_INT HIP_MatLocalityOrder(GCGE_INT nrows, const GCGE_INT *rowptr, const GCGE_INT *col, GCGE_INT leaf,
                              GCGE_INT *perm);
/* Upload P A P^T for the row order perm (perm[new position] = the caller's
 * row; NULL: HIP_MatLocalityOrder with leaf 0).  Each row keeps its entries'
 * order, so products are bit-identical row by row to A's.  Multivectors
 * created from the matrix hold rows in this order: HIP_MultiVecSetFromHost,
 * HIP_MultiVecGetToHost and the random initial vectors use the caller's
 * order, HIP_MatSetValues and the gradient Jacobians the caller's entry
 * order; A and B of one solver must share the order (error 171 otherwise).
 * MultiVecPrint prints the stored order. */
HIP_MAT *HIP_MatCreateFromCSROrdered(GCGE_INT nrows, GCGE_INT ncols, GCGE_INT nnz, const GCGE_INT *rowptr,
                                     const GCGE_INT *col, const GCGE_DOUBLE *val, const GCGE_INT *perm);
/* Writes the matrix's row order into perm[n] and returns 1, or returns 0
 * (natural order, perm untouched); -1 for a bad matrix. */
GCGE_INT HIP_MatGetOrder(const HIP_MAT *mat, GCGE_INT *perm);
/* The SpMM layout chosen at creation: 0 direct column gather (k_spmm),
 * 1 row-major gather (k_spmm_rows), 2 staged row blocks (k_spmm_halo);
 * -1 for a bad matrix.  No counterpart in the reference (a query of this
 * backend's layout choice). */
GCGE_INT HIP_MatGatherKind(const HIP_MAT *mat);
/* Write CSR arrays (host) in the text format (binary = 0; values with 17
 * significant digits, so they read back exactly) or the binary one. */
GCGE_INT HIP_WriteMatFile(const char *filename, GCGE_INT nrows, GCGE_INT ncols, GCGE_INT nnz,
                          const GCGE_INT *rowptr, const GCGE_INT *col, const GCGE_DOUBLE *val,
                          GCGE_INT binary);
void HIP_MatDestroy(HIP_MAT **mat);
GCGE_INT HIP_MatNumRows(const HIP_MAT *mat);
GCGE_INT HIP_MatNumNonzeros(const HIP_MAT *mat);
/* Replace the values (same sparsity pattern); the FP32 copy, the Jacobi
 * diagonal and the Gershgorin bound are refreshed with them. */
GCGE_INT HIP_MatSetValues(HIP_MAT *mat, const GCGE_DOUBLE *val);
/* Storage precision of the multivectors and vectors created from this matrix
 * from now on (GCGE_HIP_FP64 or GCGE_HIP_FP32).  Matrix products with FP32
 * multivectors use the FP32 copy of the values. */
GCGE_INT HIP_MatSetPrecision(HIP_MAT *mat, GCGE_INT precision);

/* ---- multivectors ---- */

GCGE_INT HIP_MultiVecCreate(void ***mv, GCGE_INT n_vec, HIP_MAT *mat);
void HIP_MultiVecDestroy(void ***mv);
/* host (column-major, leading dimension ldh, FP64) <-> columns [start, start+ncols) */
GCGE_INT HIP_MultiVecSetFromHost(void **mv, GCGE_INT start, GCGE_INT ncols, const GCGE_DOUBLE *host,
                                 GCGE_INT ldh);
GCGE_INT HIP_MultiVecGetToHost(void **mv, GCGE_INT start, GCGE_INT ncols, GCGE_DOUBLE *host,
                               GCGE_INT ldh);
GCGE_INT HIP_MultiVecNumCols(void **mv);
GCGE_INT HIP_MultiVecPrecision(void **mv);

/* ---- registration (counterparts of gcge_app_csr.h:81-84) ---- */

void GCGE_HIP_SetOps(GCGE_OPS *ops);
void GCGE_SOLVER_SetHIPOps(GCGE_SOLVER *solver);
/* Create a solver for A x = lambda B x (B may be NULL) with nev eigenpairs,
 * parse -gcge_* options from argv, allocate eval/evec (FP64 unless
 * HIP_MatSetPrecision chose FP32) and set the solver up, as
 * GCGE_CSR_Solver_Init does (app/csr/gcge_app_csr.c:92-124). */
GCGE_SOLVER *GCGE_HIP_Solver_Init(HIP_MAT *A, HIP_MAT *B, GCGE_INT num_eigenvalues, GCGE_INT argc,
                                  char *argv[]);

/* ---- mixed precision (no counterpart in the reference) ---- */

typedef struct {
    GCGE_DOUBLE fp32_tol;     /* tolerance of the FP32 phase, same criterion as para->conv_type */
    GCGE_INT fp32_max_it;     /* outer-iteration budget of the FP32 phase                      */
    GCGE_INT fp64_max_it;     /* outer-iteration budget of the FP64 refinement                 */
    GCGE_INT stagnation_window; /* FP32 phase also stops when max_res has not dropped below
                                   stagnation_ratio x its best for this many iterations (0: off) */
    GCGE_DOUBLE stagnation_ratio;
    GCGE_INT fp32_w_orth;     /* W orthonormalization of the FP32 phase: 1 CholQR ("cholqr"),
                                 0 the caller's w_orth_type                                   */
    GCGE_INT refine_w_orth;   /* the same choice for the FP64 refinement                      */
    GCGE_INT refine_bcg_precision; /* 1: block CG of the refinement in FP32 as a correction
                                      of the FP64 residual (bcg_precision), 0: FP64           */
    GCGE_DOUBLE stagnation_start;  /* stagnation is watched once max_res has fallen to this
                                      fraction of its peak (0: from the first iteration)      */
    GCGE_INT cg_max_it;       /* CG step cap of the block-CG inner solves in both phases
                                 (0: the caller's para->cg_max_it)                            */
    GCGE_INT bcg_precond;     /* 1: precondition the inner solves of both phases with the
                                 diagonal of A (Jacobi); 0: the caller's para->bcg_precond  */
    GCGE_INT rr_eig_full;     /* 1: Rayleigh-Ritz pairs from the full dsyevd decomposition in
                                 both phases (para->rr_eig_full); 0: the caller's setting     */
} GCGE_HIP_MIXED_OPTS;

typedef struct {
    GCGE_INT fp32_iters;
    GCGE_INT fp64_iters;
    GCGE_DOUBLE fp32_seconds;
    GCGE_DOUBLE fp64_seconds;
    GCGE_INT fp64_num_unlock; /* pairs above the tolerance at the final check of every pair */
    GCGE_INT refine_passes;   /* FP64 refinement passes run (restarts + 1)                 */
    GCGE_INT fp32_failed;     /* 0: the FP32 phase started the refinement; 1: it ended with
                                 a non-finite residual and was discarded; 2: skipped, A or B
                                 out of FP32 range                                         */
} GCGE_HIP_MIXED_INFO;

/* Return value of GCGE_HIP_SOLVER_SolveMixed when no error is latched but
 * pairs are still above the tolerance after the refinement budget. */
#define GCGE_HIP_NOT_CONVERGED (-2)

/* Solve with an FP32 phase (matrix values and multivectors in FP32, dense
 * Rayleigh-Ritz, Cholesky factors and the residual check in FP64) followed by
 * FP64 refinement sweeps of the same GCG iteration started from the FP32
 * eigenvectors (given_init_evec).  solver must have been created by
 * GCGE_HIP_Solver_Init with FP64 evec; its para (nev, ev_tol, conv_type, ...)
 * applies to the refinement.  When given_init_evec is set, evec holds the
 * start block.  opts may be NULL: fp32_tol 8e-3, the caller's ev_max_it for
 * both phases, stagnation window 2 / ratio 0.5 / start 1e-3, CholQR in both
 * phases, FP32 inner solves in the refinement, inner solves preconditioned
 * with the diagonal of A and capped at 30 CG steps, and Rayleigh-Ritz pairs
 * from the full dsyevd decomposition.  In opts, a zero fp32_tol, budget or
 * stagnation_ratio selects the same default, and a zero cg_max_it,
 * bcg_precond or rr_eig_full keeps the caller's.
 * Every pair is checked under the caller's criterion at the end; para->res,
 * para->max_res and para->num_unlock hold that check.  Returns 0 when every
 * pair is within ev_tol, GCGE_HIP_NOT_CONVERGED when some are not, and the
 * latched error code otherwise (a non-finite final residual latches 190).
 * The FP32 phase is skipped when the Gershgorin bound of A or B lies outside
 * [1e-18, 1e18], and discarded when it ends with a non-finite residual; the
 * FP64 iteration then starts from the caller's block (or a random one). */
GCGE_INT GCGE_HIP_SOLVER_SolveMixed(GCGE_SOLVER *solver, const GCGE_HIP_MIXED_OPTS *opts,
                                    GCGE_HIP_MIXED_INFO *info);

/* ---- eigenvalue gradients (adjoint; no counterpart in the reference) ---- */

/* For a matrix family A(theta) with a fixed pattern, let J be the sparse
 * Jacobian of A's stored values with respect to theta, given by parameter:
 * jrowptr[nparam+1], jnz[] (index of a stored nonzero of A), jval[]
 * (d A_k / d theta_p).  For the pairs (lambda_i, X(:, i)), i < k, with
 * weights w_i and B = I, writes grad[p] = sum_i w_i x_i^T (dA/dtheta_p) x_i,
 * i.e. d(sum_i w_i lambda_i)/d theta for B-normalized eigenvectors.  The
 * per-nonzero sampled outer product and the reduction onto theta run on the
 * device; grad is a host array of nparam entries. */
GCGE_INT GCGE_HIP_EigGradient(HIP_MAT *A, void **X, GCGE_INT k, const GCGE_DOUBLE *w,
                              GCGE_INT nparam, const GCGE_INT *jrowptr, const GCGE_INT *jnz,
                              const GCGE_DOUBLE *jval, GCGE_DOUBLE *grad);

/* The generalized problem A(theta) x = lambda B(theta) x: with B-normalized
 * eigenvectors X(:, i) (x_i^T B x_i = 1, as the solver returns them) and
 * their eigenvalues lambda[i], writes
 *   grad[p] = sum_i w_i x_i^T (dA/dtheta_p - lambda_i dB/dtheta_p) x_i,
 * i.e. d(sum_i w_i lambda_i)/d theta.  The A and B Jacobians are given as
 * for GCGE_HIP_EigGradient, each over its own matrix's stored entries and
 * with the same nparam rows; B == NULL (B = I) or jb_rowptr == NULL (B does
 * not depend on theta) drops the second term. */
GCGE_INT GCGE_HIP_EigGradientGeneralized(HIP_MAT *A, HIP_MAT *B, void **X, GCGE_INT k,
                                         const GCGE_DOUBLE *lambda, const GCGE_DOUBLE *w,
                                         GCGE_INT nparam, const GCGE_INT *ja_rowptr,
                                         const GCGE_INT *ja_nz, const GCGE_DOUBLE *ja_val,
                                         const GCGE_INT *jb_rowptr, const GCGE_INT *jb_nz,
                                         const GCGE_DOUBLE *jb_val, GCGE_DOUBLE *grad);

/* ---- no allocation (no counterpart in the reference) ---- */

/* For callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * MultiVecInnerProdDiag slot, out[j] = <X_{start[0]+j}, Y_{start[1]+j}>
 * (j < end[0] - start[0]), FP64 accumulation, results in host memory out,
 * with the same kernels and summation order (the same bits) and no
 * allocation.  The partial sums use a fixed device array of the library
 * (263 KiB per GPU, held until the process ends), so at most 128 columns
 * per call (larger calls are rejected with error 202); calls that overlap
 * on one GPU race on that array.  Returns 0 or the latched error code. */
GCGE_INT GCGE_HIP_MultiVecInnerProdDiagNoAlloc(void **x, void **y, GCGE_DOUBLE *out, GCGE_INT *start,
                                               GCGE_INT *end);

/* ---- errors, synchronisation, counters ---- */

/* Returns the latched error code (0 = none) and copies its message. */
GCGE_INT GCGE_HIP_GetLastError(char *msg, GCGE_INT len);
void GCGE_HIP_ClearError(void);
GCGE_INT GCGE_HIP_Synchronize(void);
/* Device memory currently held by the backend, in bytes. */
double GCGE_HIP_DeviceBytes(void);
/* Kernel profiler: when on, every launch of a kernel class is bracketed by
 * HIP events on the backend's stream and its algorithmic bytes/flops are
 * recorded.  GCGE_HIP_GetKernelStats(kind, ...) returns the number of kernel
 * classes (or -1 for a bad kind) and fills the class name, launch count,
 * summed device milliseconds, algorithmic bytes and flops. */
void GCGE_HIP_SetProfiling(GCGE_INT on);
void GCGE_HIP_ResetKernelStats(void);
GCGE_INT GCGE_HIP_GetKernelStats(GCGE_INT kind, char *name, GCGE_INT len, double *count, double *ms,
                                 double *bytes, double *flops);
/* Struct layout of this ABI for binding checks: writes up to n values
 * (sizeof GCGE_PARA and selected member offsets, sizeof GCGE_ORTH_PARA,
 * GCGE_WORKSPACE, GCGE_SOLVER, GCGE_OPS, ...) and returns how many. */
GCGE_INT GCGE_HIP_AbiLayout(GCGE_INT *out, GCGE_INT n);

#ifdef __cplusplus
}
#endif
#endif
