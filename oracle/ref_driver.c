/*
 * ref_driver.c -- runs the real GCGE-1.0 reference (compiled from its own
 * sources into oracle/_ref/libgcge_ref.so by oracle/Makefile) on a problem
 * file and writes eigenvalues, residuals, iteration counts, eigenvectors and
 * the wall time of GCGE_SOLVER_Solve.
 *
 * TEST INFRASTRUCTURE ONLY: it validates the C restatement in
 * gcge_oracle.c and serves as the "reference" CPU baseline.
 *
 * Usage: ref_solve <problem.bin> <out.bin> [-seed S] [-dirichlet 1]
 *                  [-cg_type T] [-cg_rate R] [-gcge_<flag> <value> ...]
 * The -gcge_* flags go to the reference's own parser
 * (GCGE_PARA_SetFromCommandLine, src/gcge_para.c:114-350) through
 * GCGE_CSR_Solver_Init (app/csr/gcge_app_csr.c:92-124).  When the problem
 * file carries X0, -gcge_given_init_evec 1 is implied.  cg_type has no
 * command-line flag and -gcge_cg_rate is parsed with atoi
 * (src/gcge_para.c:253), so -cg_type / -cg_rate set those fields directly
 * after initialisation, as test/csr/test_csr_solver.c:48-50 does.
 *
 * Problem file (little endian): "GCGEPRB1", int32 n, nnzA, hasB, nnzB, nev,
 * hasX0, then rowptrA[n+1] (int32), colA[nnzA] (int32), valA[nnzA] (f64),
 * the same three arrays for B when hasB, and X0[n*nev] (f64, column-major)
 * when hasX0.
 * Output file: "GCGEOUT1", int32 n, nev, num_iter, num_unlock, f64 seconds,
 * eval[nev], res[nev], evec[n*nev].
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "gcge.h"
#include "gcge_app_csr.h"

static void die(const char *m)
{
    fprintf(stderr, "ref_solve: %s\n", m);
    exit(2);
}

static void rd(void *p, size_t sz, size_t cnt, FILE *f)
{
    if (fread(p, sz, cnt, f) != cnt) die("short read");
}

static CSR_MAT *read_csr(FILE *f, int n, int nnz)
{
    CSR_MAT *M = (CSR_MAT *)malloc(sizeof(CSR_MAT));
    M->N_Rows = n;
    M->N_Columns = n;
    M->N_Entries = nnz;
    M->RowPtr = (int *)malloc((size_t)(n + 1) * sizeof(int));
    M->KCol = (int *)malloc((size_t)nnz * sizeof(int));
    M->Entries = (double *)malloc((size_t)nnz * sizeof(double));
    rd(M->RowPtr, sizeof(int), (size_t)n + 1, f);
    rd(M->KCol, sizeof(int), (size_t)nnz, f);
    rd(M->Entries, sizeof(double), (size_t)nnz, f);
    return M;
}

int main(int argc, char **argv)
{
    char magic[8];
    int hdr[6], n, nnzA, hasB, nnzB, nev, hasX0, i, seed = 1, dirichlet = 0, nargs = 0;
    int cg_type = -1;
    double cg_rate = -1.0;
    char **args;
    FILE *f;
    CSR_MAT *A, *B = NULL;
    double *X0 = NULL;
    GCGE_SOLVER *solver;
    struct timespec t0, t1;
    double secs;

    if (argc < 3) die("usage: ref_solve problem.bin out.bin [flags]");
    f = fopen(argv[1], "rb");
    if (!f) die("cannot open problem");
    rd(magic, 1, 8, f);
    if (memcmp(magic, "GCGEPRB1", 8) != 0) die("bad magic");
    rd(hdr, sizeof(int), 6, f);
    n = hdr[0]; nnzA = hdr[1]; hasB = hdr[2]; nnzB = hdr[3]; nev = hdr[4]; hasX0 = hdr[5];
    A = read_csr(f, n, nnzA);
    if (hasB) B = read_csr(f, n, nnzB);
    if (hasX0) {
        X0 = (double *)malloc((size_t)n * nev * sizeof(double));
        rd(X0, sizeof(double), (size_t)n * nev, f);
    }
    fclose(f);

    /* driver flags are consumed here, -gcge_* flags are forwarded */
    args = (char **)malloc((size_t)(argc + 4) * sizeof(char *));
    args[nargs++] = argv[0];
    for (i = 3; i < argc; i++) {
        if (strcmp(argv[i], "-seed") == 0 && i + 1 < argc) {
            seed = atoi(argv[++i]);
        } else if (strcmp(argv[i], "-dirichlet") == 0 && i + 1 < argc) {
            dirichlet = atoi(argv[++i]);
        } else if (strcmp(argv[i], "-cg_type") == 0 && i + 1 < argc) {
            cg_type = atoi(argv[++i]);
        } else if (strcmp(argv[i], "-cg_rate") == 0 && i + 1 < argc) {
            cg_rate = atof(argv[++i]);
        } else {
            args[nargs++] = argv[i];
        }
    }
    if (hasX0) {
        args[nargs++] = "-gcge_given_init_evec";
        args[nargs++] = "1";
    }
    srand((unsigned)seed);
    solver = GCGE_CSR_Solver_Init(A, B, nev, nargs, args);
    solver->para->dirichlet_boundary = dirichlet;
    if (cg_type >= 0) solver->para->cg_type = cg_type;
    if (cg_rate >= 0.0) solver->para->cg_rate = cg_rate;
    if (hasX0) {
        CSR_VEC **ev = (CSR_VEC **)solver->evec;
        for (i = 0; i < nev; i++)
            memcpy(ev[i]->Entries, X0 + (size_t)i * n, (size_t)n * sizeof(double));
    }
    fflush(stdout);
    clock_gettime(CLOCK_MONOTONIC, &t0);
    GCGE_SOLVER_Solve(solver);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    secs = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    fflush(stdout);

    f = fopen(argv[2], "wb");
    if (!f) die("cannot open output");
    fwrite("GCGEOUT1", 1, 8, f);
    {
        int oh[4];
        oh[0] = n; oh[1] = nev; oh[2] = solver->para->num_iter; oh[3] = solver->para->num_unlock;
        fwrite(oh, sizeof(int), 4, f);
    }
    fwrite(&secs, sizeof(double), 1, f);
    fwrite(solver->eval, sizeof(double), (size_t)nev, f);
    fwrite(solver->para->res, sizeof(double), (size_t)nev, f);
    {
        CSR_VEC **ev = (CSR_VEC **)solver->evec;
        for (i = 0; i < nev; i++) fwrite(ev[i]->Entries, sizeof(double), (size_t)n, f);
    }
    fclose(f);
    return 0;
}
