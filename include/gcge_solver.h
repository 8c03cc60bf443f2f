/*
 * gcge_solver.h -- workspace, solver object and solver API.
 *
 * GCGE_WORKSPACE and GCGE_SOLVER keep the member layout of the reference
 * (src/gcge_workspace.h:26-66, src/gcge_solver.h:28-39); the GCGE_SOLVER_*
 * functions have the reference's names and signatures
 * (src/gcge_solver.h:42-57) and GCGE_Solve / GCGE_CheckConvergence /
 * GCGE_CheckEvalMultiplicity those of src/gcge_eigsol.h:48-52.
 *
 * Ownership follows the reference: the caller owns A, B, eval and evec
 * (GCGE_SOLVER_Free_All destroys evec through ops->MultiVecDestroy and frees
 * eval); the solver owns V, V_tmp and CG_p.  The solver uses evec as scratch
 * during the iteration, as the reference does (gcge_rayleighritz.c:198-199,
 * gcge_xpw.c:199-201).
 */
#ifndef GCGE_SOLVER_H_
#define GCGE_SOLVER_H_

#include "gcge_config.h"
#include "gcge_ops.h"
#include "gcge_para.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GCGE_WORKSPACE_ {
    GCGE_INT V_size;
    GCGE_INT V_tmp_size;
    GCGE_INT CG_p_size;
    GCGE_INT CG_r_size;
    GCGE_INT RitzVec_size;
    void **V;        /* basis [X, P, W]                    */
    void **V_tmp;    /* scratch columns                    */
    void **RitzVec;
    void **evec;     /* the caller's eigenvector block     */
    void **CG_p;     /* CG search directions               */
    void **CG_r;

    GCGE_INT *unlock;            /* indices of unconverged pairs */
    GCGE_INT max_dim_x;
    GCGE_INT dim_x;
    GCGE_INT dim_xp;
    GCGE_INT dim_xpw;
    GCGE_INT last_dim_xpw;
    GCGE_INT last_dim_x;

    GCGE_INT conv_bs;
    GCGE_INT unconv_bs;
    GCGE_INT num_soft_locked_in_last_iter;

    void *dense_matrix;

    GCGE_DOUBLE *subspace_matrix;
    GCGE_DOUBLE *subspace_evec;
    GCGE_DOUBLE *subspace_dtmp;
    GCGE_INT *subspace_itmp;

    GCGE_DOUBLE *eval;

    GCGE_INT *orth_ind;

    /* ---- members of this build (no counterpart in the reference) ---- */
    void **bcg32[4];     /* FP32 block-CG work columns (rhs/residual, solution, p, w) */
    GCGE_INT bcg32_cols; /* columns allocated in each of them                        */
    void *mixed32;       /* solver of the FP32 phase of GCGE_HIP_SOLVER_SolveMixed,
                            kept for later calls (NULL until the first)              */
} GCGE_WORKSPACE;

typedef struct GCGE_SOLVER_ {
    void *A;
    void *B;
    void **evec;
    GCGE_DOUBLE *eval;
    GCGE_PARA *para;
    GCGE_OPS *ops;
    GCGE_WORKSPACE *workspace;
} GCGE_SOLVER;

void GCGE_WORKSPACE_Create(GCGE_WORKSPACE **workspace);
void GCGE_WORKSPACE_Setup(GCGE_WORKSPACE *workspace, GCGE_PARA *para, GCGE_OPS *ops, void *A);
void GCGE_WORKSPACE_Free(GCGE_WORKSPACE **workspace, GCGE_PARA *para, GCGE_OPS *ops);

void GCGE_SOLVER_Create(GCGE_SOLVER **solver);
void GCGE_SOLVER_Free(GCGE_SOLVER **solver);
void GCGE_SOLVER_Free_All(GCGE_SOLVER **solver);
void GCGE_SOLVER_SetMatA(GCGE_SOLVER *solver, void *A);
void GCGE_SOLVER_SetMatB(GCGE_SOLVER *solver, void *B);
void GCGE_SOLVER_SetEigenvalues(GCGE_SOLVER *solver, GCGE_DOUBLE *eval);
void GCGE_SOLVER_SetEigenvectors(GCGE_SOLVER *solver, void **evec);
void GCGE_SOLVER_SetNumEigen(GCGE_SOLVER *solver, GCGE_INT nev);
void GCGE_SOLVER_Setup(GCGE_SOLVER *solver);
void GCGE_SOLVER_SetOpsLinearSolverWorkspace(GCGE_SOLVER *solver, void *linear_solver_workspace);
void GCGE_SOLVER_GetEigenvalues(GCGE_SOLVER *solver, GCGE_DOUBLE **eval);
void GCGE_SOLVER_GetEigenvectors(GCGE_SOLVER *solver, void ***evec);
void GCGE_SOLVER_Solve(GCGE_SOLVER *solver);

void GCGE_Solve(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_PARA *para, GCGE_OPS *ops,
                GCGE_WORKSPACE *workspace);
void GCGE_CheckConvergence(void *A, void *B, GCGE_DOUBLE *eval, void **evec, GCGE_OPS *ops,
                           GCGE_PARA *para, GCGE_WORKSPACE *workspace);
void GCGE_CheckEvalMultiplicity(GCGE_INT start, GCGE_INT nev, GCGE_INT end, GCGE_INT *dim_x,
                                GCGE_DOUBLE *eval);

#ifdef __cplusplus
}
#endif
#endif
