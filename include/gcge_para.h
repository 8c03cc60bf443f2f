/*
 * gcge_para.h -- solver parameters and statistics.
 *
 * GCGE_ORTH_PARA, GCGE_PARA, GCGE_PART_TIME, GCGE_OPS_TIME_COUNT and
 * GCGE_STATISTIC_PARA keep the member names, types and order of the
 * reference (src/gcge_para.h:26-103, src/gcge_statistics.h:23-49), because
 * applications set parameters by writing the struct members directly
 * (e.g. example/csr_solver.c:51, test/csr/test_csr_solver.c:46-52).
 * Members of this build are appended after the reference's last member.
 */
#ifndef GCGE_PARA_H_
#define GCGE_PARA_H_

#include "gcge_config.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GCGE_PART_TIME_ {
    GCGE_DOUBLE x_axpy_time;
    GCGE_DOUBLE x_orth_time;
    GCGE_DOUBLE p_axpy_time;
    GCGE_DOUBLE p_orth_time;
    GCGE_DOUBLE w_line_time;   /* linear solves that build W */
    GCGE_DOUBLE w_orth_time;
    GCGE_DOUBLE rr_mat_time;   /* Rayleigh-Ritz subspace matrix */
    GCGE_DOUBLE rr_eigen_time; /* dense eigensolve */
    GCGE_DOUBLE conv_time;
} GCGE_PART_TIME;

typedef struct GCGE_OPS_TIME_COUNT_ {
    GCGE_DOUBLE mv_time;
    GCGE_INT mv_count;
    GCGE_DOUBLE ip_time;
    GCGE_INT ip_count;
    GCGE_DOUBLE axpy_time;
    GCGE_INT axpy_count;
} GCGE_OPS_TIME_COUNT;

typedef struct GCGE_STATISTIC_PARA_ {
    GCGE_PART_TIME *part_time_total;
    GCGE_PART_TIME *part_time_one_iter;
    GCGE_OPS_TIME_COUNT *ops_time_count_total;
    GCGE_OPS_TIME_COUNT *ops_time_count_linear_solver;
} GCGE_STATISTIC_PARA;

typedef struct GCGE_ORTH_PARA_ {
    GCGE_DOUBLE orth_zero_tol;    /* a vector with norm below this is dropped       */
    GCGE_DOUBLE reorth_tol;       /* reorthogonalize while the norm ratio is below  */
    GCGE_DOUBLE scbgs_reorth_tol; /* stop block reorthogonalization below this norm */
    GCGE_DOUBLE criterion_tol;    /* column-wise Gram-Schmidt once sum_res is below */
    GCGE_INT max_reorth_time;
    GCGE_INT scbgs_wself_max_reorth_time;
    GCGE_INT print_orth_zero;
    GCGE_INT x_orth_block_size;
    GCGE_INT w_orth_block_size;
    GCGE_INT max_direct_orth_length; /* leaf size of the recursive orthogonalization */
} GCGE_ORTH_PARA;

typedef struct GCGE_PARA_ {
    GCGE_INT nev;
    GCGE_INT ev_max_it;
    GCGE_INT block_size;

    GCGE_INT if_lobgcg;
    GCGE_INT given_init_evec;

    GCGE_DOUBLE ev_tol;
    char *conv_type;             /* "R" relative, "O" omega, "A" absolute, "NS" (this build) */
    GCGE_DOUBLE conv_omega_norm;
    GCGE_INT num_unlock;

    char *orth_type;
    char *w_orth_type;
    char *p_orth_type;
    char *x_orth_type;
    GCGE_ORTH_PARA *orth_para;

    GCGE_DOUBLE multi_tol;
    GCGE_DOUBLE multi_tol_for_lock;

    GCGE_INT if_use_cg;
    GCGE_INT cg_max_it;
    GCGE_DOUBLE cg_rate;
    GCGE_INT cg_type;
    GCGE_INT if_use_bcg;

    GCGE_INT num_iter;
    GCGE_DOUBLE max_res;
    GCGE_DOUBLE max_ind;
    GCGE_DOUBLE min_res;
    GCGE_DOUBLE min_ind;
    GCGE_DOUBLE sum_res;
    GCGE_DOUBLE *res;
    GCGE_INT dirichlet_boundary;

    GCGE_INT use_mpi_bcast;

    GCGE_INT print_eval;
    GCGE_INT print_cg_error;
    GCGE_INT print_orthzero;
    GCGE_INT print_para;
    GCGE_INT print_result;
    GCGE_INT print_matlab;
    GCGE_INT print_part_time;
    GCGE_INT print_level;

    GCGE_STATISTIC_PARA *stat_para;

    GCGE_INT opt_rr_eig_partly;
    GCGE_INT opt_bcast;
    GCGE_INT opt_allgatherv;

    /* ---- members of this build (no counterpart in the reference) ---- */
    GCGE_INT bcg_iters_total;    /* output: CG sweeps summed over all BCG calls   */
    GCGE_INT bcg_calls;          /* output: number of BCG calls                   */
    GCGE_INT spmv_columns;       /* output: matrix-vector columns applied (A, B)  */
    GCGE_DOUBLE solve_seconds;   /* output: wall time of the last GCGE_Solve      */
    /* input: when > 0, the outer iteration also stops once max_res has not
     * fallen below stagnation_ratio times its best value for this many
     * iterations, or at once when it rises above 10 times its lowest value
     * (a basis that has lost its orthonormality gives spurious Ritz
     * values).  While watched, the iterate with the lowest max_res is kept
     * in a copy and restored when the iteration stops without converging
     * (used by the FP32 phase of the mixed-precision solve) */
    GCGE_INT stagnation_window;
    GCGE_DOUBLE stagnation_ratio;
    GCGE_INT stopped_on_stagnation; /* output: 1 stagnation, 2 breakdown */
    /* input: storage precision of the block-CG inner solves that build W
     * (0: same as the basis; 1: FP32 defect correction -- the initial
     * residual r0 = b - A W0 is formed in FP64, the CG for the correction
     * A E = r0 runs on FP32 vectors and FP32 matrix values with
     * FP64-accumulated dots, and the W block is E, widened to FP64: W0 is a
     * scaled column of X, so W0 + E and E span the same space with X).
     * Needs the MultiVecCreateByMatPrecision slot. */
    GCGE_INT bcg_precision;
    /* input: when > 0, stagnation is only watched once max_res has fallen to
     * this fraction of its largest value so far (residuals of a random start
     * first grow before they fall) */
    GCGE_DOUBLE stagnation_start;
    /* input: 1 (the reference's behaviour) removes converged pairs from the
     * W and P blocks; 0 keeps every pair active until all have converged
     * (num_unlock stays nev until then) */
    GCGE_INT lock_converged;
    /* input: unit roundoff of the basis storage (DBL_EPSILON for FP64,
     * FLT_EPSILON for FP32); sets the drop and check tolerances of the
     * "cholqr" W orthonormalization */
    GCGE_DOUBLE basis_eps;
    /* input: 1 preconditions the block CG that builds W with the diagonal of
     * A (Jacobi); 0, the default, is the reference's unpreconditioned BCG.
     * The stopping rule stays ||r_k|| / ||r_0|| < cg_rate.  Needs the
     * BlockPCGResidualUpdate slot. */
    GCGE_INT bcg_precond;
    /* input: 1 takes the Rayleigh-Ritz eigenpairs il..iu from the full
     * decomposition of the subspace matrix (LAPACK dsyevd, divide and
     * conquer: about 3x faster than dsyevx at 300 x 300, and orthonormal
     * eigenvectors inside clusters of equal eigenvalues) instead of the
     * DenseMatEigenSolver slot (dsyevx with range "I", the reference's
     * choice, kept by default: iteration counts follow it). */
    GCGE_INT rr_eig_full;
} GCGE_PARA;

void GCGE_PARA_Create(GCGE_PARA **para);
void GCGE_PARA_Free(GCGE_PARA **para);
GCGE_INT GCGE_PARA_SetFromCommandLine(GCGE_PARA *para, GCGE_INT argc, char **argv);
void GCGE_PARA_SetNumEigen(GCGE_PARA *para, GCGE_INT nev);
void GCGE_PARA_Setup(GCGE_PARA *para);
void GCGE_PrintParaInfo(GCGE_PARA *para);
void GCGE_PrintIterationInfo(GCGE_DOUBLE *eval, GCGE_PARA *para);
void GCGE_PrintFinalInfo(GCGE_DOUBLE *eval, GCGE_PARA *para);

void GCGE_STATISTIC_PARA_Create(GCGE_STATISTIC_PARA **stat_para);
void GCGE_STATISTIC_PARA_Free(GCGE_STATISTIC_PARA **stat_para);

#ifdef __cplusplus
}
#endif
#endif
