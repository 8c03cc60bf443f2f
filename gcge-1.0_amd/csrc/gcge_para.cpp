// gcge_para.cpp -- parameters, statistics and reports of the GCGE solver.
//
// Defaults, command-line flags and GCGE_PARA_Setup follow the reference
// (src/gcge_para.c:25-99, :114-350, :357-386).  One deliberate difference:
// -gcge_cg_rate is parsed as a floating-point number (the reference uses
// atoi, src/gcge_para.c:253, which turns every rate below 1 into 0).
#include <cmath>
#include <cfloat>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "gcge.h"

extern "C" void GCGE_STATISTIC_PARA_Create(GCGE_STATISTIC_PARA **stat_para)
{
    GCGE_STATISTIC_PARA *s = (GCGE_STATISTIC_PARA *)calloc(1, sizeof(GCGE_STATISTIC_PARA));
    s->part_time_total = (GCGE_PART_TIME *)calloc(1, sizeof(GCGE_PART_TIME));
    s->part_time_one_iter = (GCGE_PART_TIME *)calloc(1, sizeof(GCGE_PART_TIME));
    s->ops_time_count_total = (GCGE_OPS_TIME_COUNT *)calloc(1, sizeof(GCGE_OPS_TIME_COUNT));
    s->ops_time_count_linear_solver = (GCGE_OPS_TIME_COUNT *)calloc(1, sizeof(GCGE_OPS_TIME_COUNT));
    *stat_para = s;
}

extern "C" void GCGE_STATISTIC_PARA_Free(GCGE_STATISTIC_PARA **stat_para)
{
    if (!stat_para || !*stat_para) return;
    free((*stat_para)->part_time_total);
    free((*stat_para)->part_time_one_iter);
    free((*stat_para)->ops_time_count_total);
    free((*stat_para)->ops_time_count_linear_solver);
    free(*stat_para);
    *stat_para = nullptr;
}

extern "C" void GCGE_PARA_Create(GCGE_PARA **para)
{
    GCGE_PARA *p = (GCGE_PARA *)calloc(1, sizeof(GCGE_PARA));
    p->nev = 1;
    p->ev_max_it = 30;
    p->block_size = 0;
    p->if_lobgcg = 0;
    p->given_init_evec = 0;
    p->ev_tol = 1e-8;
    p->conv_type = (char *)"R";
    p->orth_type = (char *)"B";
    p->w_orth_type = (char *)"multi";
    p->p_orth_type = (char *)"multi";
    p->x_orth_type = (char *)"multi";
    p->conv_omega_norm = 0.0;
    p->orth_para = (GCGE_ORTH_PARA *)calloc(1, sizeof(GCGE_ORTH_PARA));
    p->orth_para->orth_zero_tol = 1e-16;
    p->orth_para->reorth_tol = 0.75;
    p->orth_para->scbgs_reorth_tol = 1e-16;
    p->orth_para->max_reorth_time = 3;
    p->orth_para->scbgs_wself_max_reorth_time = 1;
    p->orth_para->criterion_tol = 1e-6;
    p->orth_para->print_orth_zero = 0;
    p->orth_para->x_orth_block_size = 0;
    p->orth_para->w_orth_block_size = 0;
    p->orth_para->max_direct_orth_length = 1;
    p->multi_tol = 0.2;
    p->multi_tol_for_lock = 1e-6;
    p->if_use_cg = 1;
    p->cg_max_it = 20;
    p->cg_rate = 1e-2;
    p->cg_type = 1;
    p->if_use_bcg = 1;
    p->num_iter = 0;
    p->num_unlock = 6;
    p->use_mpi_bcast = 1;
    p->print_cg_error = 0;
    p->print_eval = 0;
    p->print_part_time = 0;
    p->print_para = 1;
    p->print_result = 1;
    p->max_res = 0.0;
    p->max_ind = 0;
    p->min_res = 0.0;
    p->min_ind = 0;
    p->sum_res = 0.0;
    p->res = nullptr;
    p->dirichlet_boundary = 0;
    GCGE_STATISTIC_PARA_Create(&p->stat_para);
    p->opt_rr_eig_partly = 0;
    p->opt_bcast = 1;
    p->opt_allgatherv = 0;
    p->stagnation_start = 0.0;
    p->lock_converged = 1;
    p->basis_eps = DBL_EPSILON;
    p->bcg_precond = 0;
    p->rr_eig_full = 0;
    *para = p;
}

extern "C" void GCGE_PARA_Free(GCGE_PARA **para)
{
    if (!para || !*para) return;
    GCGE_STATISTIC_PARA_Free(&(*para)->stat_para);
    free((*para)->orth_para);
    free((*para)->res);
    free(*para);
    *para = nullptr;
}

extern "C" GCGE_INT GCGE_PARA_SetFromCommandLine(GCGE_PARA *p, GCGE_INT argc, char **argv)
{
    int i = 0, usage = 0;
    GCGE_ORTH_PARA *o = p->orth_para;
    while (i < argc) {
        const char *a = argv[i];
        const char *v = (i + 1 < argc) ? argv[i + 1] : nullptr;
        auto is = [&](const char *f) { return std::strcmp(a, f) == 0 && v; };
        if (is("-gcge_nev")) p->nev = atoi(v);
        else if (is("-gcge_ev_max_it")) p->ev_max_it = atoi(v);
        else if (is("-gcge_block_size")) p->block_size = atoi(v);
        else if (is("-gcge_x_orth_block_size")) o->x_orth_block_size = atoi(v);
        else if (is("-gcge_w_orth_block_size")) o->w_orth_block_size = atoi(v);
        else if (is("-gcge_max_direct_orth_length")) o->max_direct_orth_length = atoi(v);
        else if (is("-gcge_if_lobgcg")) p->if_lobgcg = atoi(v);
        else if (is("-gcge_given_init_evec")) p->given_init_evec = atoi(v);
        else if (is("-gcge_ev_tol")) p->ev_tol = atof(v);
        else if (is("-gcge_conv_type")) p->conv_type = argv[i + 1];
        else if (is("-gcge_conv_omega_norm")) p->conv_omega_norm = atof(v);
        else if (is("-gcge_orth_type")) p->orth_type = argv[i + 1];
        else if (is("-gcge_w_orth_type")) p->w_orth_type = argv[i + 1];
        else if (is("-gcge_p_orth_type")) p->p_orth_type = argv[i + 1];
        else if (is("-gcge_x_orth_type")) p->x_orth_type = argv[i + 1];
        else if (is("-gcge_orth_zero_tol")) o->orth_zero_tol = atof(v);
        else if (is("-gcge_reorth_tol")) o->reorth_tol = atof(v);
        else if (is("-gcge_scbgs_reorth_tol")) o->scbgs_reorth_tol = atof(v);
        else if (is("-gcge_max_reorth_time")) o->max_reorth_time = atoi(v);
        else if (is("-gcge_scbgs_wself_max_reorth_time")) o->scbgs_wself_max_reorth_time = atoi(v);
        else if (is("-gcge_print_orth_zero")) o->print_orth_zero = atoi(v);
        else if (is("-gcge_multi_tol")) p->multi_tol = atof(v);
        else if (is("-gcge_multi_tol_for_lock")) p->multi_tol_for_lock = atof(v);
        else if (is("-gcge_if_use_cg")) p->if_use_cg = atoi(v);
        else if (is("-gcge_if_use_bcg")) p->if_use_bcg = atoi(v);
        else if (is("-gcge_cg_max_it")) p->cg_max_it = atoi(v);
        else if (is("-gcge_cg_rate")) p->cg_rate = atof(v);
        else if (is("-gcge_cg_type")) p->cg_type = atoi(v);
        else if (is("-gcge_print_cg_error")) p->print_cg_error = atoi(v);
        else if (is("-gcge_print_eval")) p->print_eval = atoi(v);
        else if (is("-gcge_print_part_time")) p->print_part_time = atoi(v);
        else if (is("-gcge_print_para")) p->print_para = atoi(v);
        else if (is("-gcge_print_result")) p->print_result = atoi(v);
        else if (is("-gcge_opt_rr_eig_partly")) p->opt_rr_eig_partly = atoi(v);
        else if (is("-gcge_opt_bcast")) p->opt_bcast = atoi(v);
        else if (is("-gcge_opt_allgatherv")) p->opt_allgatherv = atoi(v);
        else if (is("-gcge_bcg_precision")) p->bcg_precision = atoi(v);
        else if (is("-gcge_lock_converged")) p->lock_converged = atoi(v);
        else if (is("-gcge_bcg_precond")) p->bcg_precond = atoi(v);
        else if (is("-gcge_rr_eig_full")) p->rr_eig_full = atoi(v);
        else if (std::strcmp(a, "-gcge_help") == 0) {
            usage = 1;
            break;
        } else {
            i++;
            continue;
        }
        i += 2;
    }
    if (usage) {
        printf("Usage: [-gcge_nev i] [-gcge_ev_max_it i] [-gcge_block_size i] [-gcge_given_init_evec i]\n"
               "       [-gcge_ev_tol d] [-gcge_conv_type R|O|A|NS] [-gcge_w_orth_type multi|gs|cholqr]\n"
               "       [-gcge_orth_zero_tol d] [-gcge_reorth_tol d] [-gcge_scbgs_reorth_tol d]\n"
               "       [-gcge_max_reorth_time i] [-gcge_max_direct_orth_length i] [-gcge_cg_max_it i]\n"
               "       [-gcge_cg_rate d] [-gcge_cg_type 1|2] [-gcge_if_use_bcg 0|1] [-gcge_print_* i]\n"
               "       [-gcge_bcg_precision 0|1] [-gcge_bcg_precond 0|1] [-gcge_lock_converged 0|1]\n"
               "       [-gcge_rr_eig_full 0|1]\n");
        return 1;
    }
    return 0;
}

extern "C" void GCGE_PARA_SetNumEigen(GCGE_PARA *para, GCGE_INT nev) { para->nev = nev; }

// GCGE_PARA_Setup, src/gcge_para.c:357-386
extern "C" void GCGE_PARA_Setup(GCGE_PARA *p)
{
    const int nev = p->nev;
    p->num_unlock = nev;
    if (p->block_size == 0) p->block_size = nev;
    if (p->orth_para->x_orth_block_size == 0) p->orth_para->x_orth_block_size = nev / 5;
    if (p->orth_para->w_orth_block_size == 0)
        p->orth_para->w_orth_block_size = (p->block_size / 5 > 1) ? (p->block_size / 5) : 1;
    p->orth_para->orth_zero_tol += DBL_EPSILON;
    p->orth_para->scbgs_reorth_tol += DBL_EPSILON;
    free(p->res);
    p->res = (GCGE_DOUBLE *)calloc(nev, sizeof(GCGE_DOUBLE));
    for (int i = 0; i < nev; i++) p->res[i] = 1.0;
}

extern "C" void GCGE_PrintParaInfo(GCGE_PARA *p)
{
    if (!p->print_para) return;
    printf("GCGE parameters: nev %d, ev_max_it %d, block_size %d, ev_tol %e, conv_type %s,\n"
           "  orth_type %s, w_orth_type %s, p_orth_type %s, orth_zero_tol %e, reorth_tol %e,\n"
           "  scbgs_reorth_tol %e, max_reorth_time %d, max_direct_orth_length %d,\n"
           "  cg_max_it %d, cg_rate %e, cg_type %d, if_use_bcg %d, given_init_evec %d\n",
           p->nev, p->ev_max_it, p->block_size, p->ev_tol, p->conv_type, p->orth_type, p->w_orth_type,
           p->p_orth_type, p->orth_para->orth_zero_tol, p->orth_para->reorth_tol, p->orth_para->scbgs_reorth_tol,
           p->orth_para->max_reorth_time, p->orth_para->max_direct_orth_length, p->cg_max_it, p->cg_rate,
           p->cg_type, p->if_use_bcg, p->given_init_evec);
    fflush(stdout);
}

static const char *conv_name(const char *ct)
{
    if (std::strcmp(ct, "R") == 0) return "relative_res";
    if (std::strcmp(ct, "O") == 0) return "omega_res";
    return "abosolute_res";
}

// Same report as GCGE_PrintIterationInfo, src/gcge_para.c:390-437.
extern "C" void GCGE_PrintIterationInfo(GCGE_DOUBLE *eval, GCGE_PARA *p)
{
    printf("num_unlock(%3d) = %3d; max_res(%3d) = %e; min_res(%3d) = %e;\n", p->num_iter, p->num_unlock,
           p->num_iter, p->max_res, p->num_iter, p->min_res);
    if (p->print_eval == 1) {
        const char *ct = conv_name(p->conv_type);
        for (int i = 0; i < p->nev; i++) {
            printf("eval(%d,%3d) = %20.15e ; %s(%d,%3d) =  %20.15e ;", p->num_iter, i + 1, eval[i], ct, p->num_iter,
                   i + 1, p->res[i]);
            if (p->res[i] < p->ev_tol) printf(" (converged)\n");
            else if (i == p->max_ind) printf(" (max)\n");
            else if (i == p->min_ind) printf(" (min)\n");
            else printf("\n");
        }
    }
    if (p->print_part_time == 1) {
        GCGE_PART_TIME *t = p->stat_para->part_time_one_iter;
        printf("iteration %d: x_orth %.4f x_axpy %.4f p_orth %.4f p_axpy %.4f conv %.4f w_lsol %.4f "
               "w_orth %.4f rr_eig %.4f rr_mat %.4f (s)\n",
               p->num_iter, t->x_orth_time, t->x_axpy_time, t->p_orth_time, t->p_axpy_time, t->conv_time,
               t->w_line_time, t->w_orth_time, t->rr_eigen_time, t->rr_mat_time);
    }
    fflush(stdout);
}

// Final table, as GCGE_PrintFinalInfo (src/gcge_para.c:502-560).
extern "C" void GCGE_PrintFinalInfo(GCGE_DOUBLE *eval, GCGE_PARA *p)
{
    GCGE_PART_TIME *t = p->stat_para->part_time_total;
    double total = t->x_axpy_time + t->x_orth_time + t->p_axpy_time + t->p_orth_time + t->w_line_time +
                   t->w_orth_time + t->rr_mat_time + t->rr_eigen_time + t->conv_time;
    if (p->print_result) {
        printf("\nFinal eigenvalues and residual information:\n\n            eigenvalue           ");
        if (std::strcmp(p->conv_type, "R") == 0) printf("relative residual\n\n");
        else if (std::strcmp(p->conv_type, "O") == 0) printf("omega residual\n\n");
        else printf("abosolute residual\n\n");
        for (int i = 0; i < p->nev; i++) {
            printf("%5d  %20.15e  %20.15e ", i + 1, eval[i], p->res[i]);
            if (p->res[i] < p->ev_tol) printf(" (converged)\n");
            else printf("\n");
        }
        printf("\nTotal gcg iteration: %d, total time: %.2lf. ", p->num_iter, total);
        if (p->num_unlock == 0) printf("All %d eigenpairs converged!\n", p->nev);
        else printf("%d converged, %d unconverged.\n", p->nev - p->num_unlock, p->num_unlock);
    }
    fflush(stdout);
}
