// gcge_dense.cpp -- dense FP64 slots of GCGE_OPS (host LAPACK/BLAS).
//
// The subspace problems are at most (max_dim_x + 2 block_size)^2 doubles
// (308 x 308 for 100 eigenpairs), so they stay on the host in FP64 and call
// the same routines as the reference's defaults (src/gcge_ops.c:302-429):
// dsyevx with range "I", dgemm, dsymm and BLAS-1.
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gcge.h"

// LAPACK/BLAS are resolved at run time from the library named by
// GCGE_LAPACK_LIB (default: MKL's single dynamic library), so that no
// foreign directory enters the loader search path of the HIP runtime.
namespace {

typedef void (*syevx_t)(const char *, const char *, const char *, int *, double *, int *, double *, double *, int *,
                        int *, double *, int *, double *, double *, int *, double *, int *, int *, int *, int *);
typedef void (*gemm_t)(const char *, const char *, int *, int *, int *, double *, double *, int *, double *, int *,
                       double *, double *, int *);
typedef void (*symm_t)(const char *, const char *, int *, int *, double *, double *, int *, double *, int *, double *,
                       double *, int *);
typedef double (*dot_t)(int *, double *, int *, double *, int *);
typedef double (*nrm2_t)(int *, double *, int *);
typedef void (*scal_t)(int *, double *, double *, int *);
typedef void (*axpy_t)(int *, double *, double *, int *, double *, int *);
typedef void (*copy_t)(int *, double *, int *, double *, int *);
typedef void (*potrf_t)(const char *, int *, double *, int *, int *);
typedef void (*trtri_t)(const char *, const char *, int *, double *, int *, int *);
typedef void (*syev_t)(const char *, const char *, int *, double *, int *, double *, double *, int *, int *);
typedef void (*syevd_t)(const char *, const char *, int *, double *, int *, double *, double *, int *, int *, int *,
                        int *);

struct Lapack {
    bool ok = false;
    syevx_t dsyevx_ = nullptr;
    gemm_t dgemm_ = nullptr;
    symm_t dsymm_ = nullptr;
    dot_t ddot_ = nullptr;
    nrm2_t dnrm2_ = nullptr;
    scal_t dscal_ = nullptr;
    axpy_t daxpy_ = nullptr;
    copy_t dcopy_ = nullptr;
    potrf_t dpotrf_ = nullptr;  // used by the CholQR orthonormalization only
    trtri_t dtrtri_ = nullptr;
    syev_t dsyev_ = nullptr;    // fallback of the subspace eigensolver (see d_eig)
    syevd_t dsyevd_ = nullptr;  // full decomposition for para->rr_eig_full
};

Lapack &lapack()
{
    static Lapack L;
    static bool tried = false;
    if (tried) return L;
    tried = true;
    const char *name = getenv("GCGE_LAPACK_LIB");
    if (!name || !*name) name = "/opt/conda/lib/libmkl_rt.so.1";
    void *h = dlopen(name, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "GCGE: cannot load LAPACK library %s: %s\n", name, dlerror());
        return L;
    }
    L.dsyevx_ = (syevx_t)dlsym(h, "dsyevx_");
    L.dgemm_ = (gemm_t)dlsym(h, "dgemm_");
    L.dsymm_ = (symm_t)dlsym(h, "dsymm_");
    L.ddot_ = (dot_t)dlsym(h, "ddot_");
    L.dnrm2_ = (nrm2_t)dlsym(h, "dnrm2_");
    L.dscal_ = (scal_t)dlsym(h, "dscal_");
    L.daxpy_ = (axpy_t)dlsym(h, "daxpy_");
    L.dcopy_ = (copy_t)dlsym(h, "dcopy_");
    L.dpotrf_ = (potrf_t)dlsym(h, "dpotrf_");
    L.dtrtri_ = (trtri_t)dlsym(h, "dtrtri_");
    L.dsyev_ = (syev_t)dlsym(h, "dsyev_");
    L.dsyevd_ = (syevd_t)dlsym(h, "dsyevd_");
    L.ok = L.dsyevx_ && L.dgemm_ && L.dsymm_ && L.ddot_ && L.dnrm2_ && L.dscal_ && L.daxpy_ && L.dcopy_;
    if (!L.ok) fprintf(stderr, "GCGE: LAPACK library %s lacks a required routine\n", name);
    return L;
}

#define LAPACK_CALL(fn, ...)                                                     \
    do {                                                                         \
        Lapack &L_ = lapack();                                                   \
        if (!L_.ok) {                                                            \
            fprintf(stderr, "GCGE: no LAPACK available for " #fn "\n");           \
            abort();                                                             \
        }                                                                        \
        L_.fn(__VA_ARGS__);                                                      \
    } while (0)

double lapack_dot(int *n, double *x, int *incx, double *y, int *incy)
{
    Lapack &L = lapack();
    if (!L.ok) abort();
    return L.ddot_(n, x, incx, y, incy);
}

double lapack_nrm2(int *n, double *x, int *incx)
{
    Lapack &L = lapack();
    if (!L.ok) abort();
    return L.dnrm2_(n, x, incx);
}

}  // namespace

namespace {

// dsyevx as the reference calls it (gcge_ops.c:302-336).  When dsyevx
// reports eigenvectors that failed to converge (info > 0: inverse iteration
// inside a cluster of equal eigenvalues, as when a graph has more isolated
// vertices than eigenpairs are wanted), the pairs il..iu are recomputed from
// the full decomposition of dsyev (QR iteration, orthonormal eigenvectors for
// any spectrum) of the same matrix; info then reports dsyev's outcome.
void d_eig(char *jobz, char *range, char *uplo, GCGE_INT *nrows, GCGE_DOUBLE *a, GCGE_INT *lda, GCGE_DOUBLE *vl,
           GCGE_DOUBLE *vu, GCGE_INT *il, GCGE_INT *iu, GCGE_DOUBLE *abstol, GCGE_INT *nev, GCGE_DOUBLE *eval,
           GCGE_DOUBLE *evec, GCGE_INT *lde, GCGE_INT *, GCGE_DOUBLE *work, GCGE_INT *lwork, GCGE_INT *iwork,
           GCGE_INT *, GCGE_INT *ifail, GCGE_INT *info)
{
    Lapack &L = lapack();
    const int n = *nrows, ld = *lda;
    const bool can_redo = L.dsyev_ && *jobz == 'V' && *range == 'I' && n > 0;
    std::vector<double> keep;
    // dsyevx destroys its triangle of a; the last column ends at row n
    if (can_redo) keep.assign(a, a + (size_t)ld * (n - 1) + n);
    LAPACK_CALL(dsyevx_, jobz, range, uplo, nrows, a, lda, vl, vu, il, iu, abstol, nev, eval, evec, lde, work, lwork,
                iwork, ifail, info);
    if (!can_redo || *info <= 0) return;
    std::vector<double> w(n), wk(std::max(1, 3 * n - 1));
    int lw = (int)wk.size(), inf = 0, nn = n, ldk = ld;
    keep.resize((size_t)ld * n);
    L.dsyev_("V", uplo, &nn, keep.data(), &ldk, w.data(), wk.data(), &lw, &inf);
    if (inf != 0) {
        *info = inf;
        return;
    }
    const int m = *iu - *il + 1;
    for (int j = 0; j < m; j++) {
        eval[j] = w[*il - 1 + j];
        std::memcpy(evec + (size_t)j * *lde, keep.data() + (size_t)(*il - 1 + j) * ld, sizeof(double) * n);
        ifail[j] = 0;
    }
    *nev = m;
    *info = 0;
}

void d_gemm(char *ta, char *tb, GCGE_INT *m, GCGE_INT *n, GCGE_INT *k, GCGE_DOUBLE *alpha, GCGE_DOUBLE *a,
            GCGE_INT *lda, GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta, GCGE_DOUBLE *c, GCGE_INT *ldc)
{
    LAPACK_CALL(dgemm_, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

void d_symm(char *side, char *uplo, GCGE_INT *m, GCGE_INT *n, GCGE_DOUBLE *alpha, GCGE_DOUBLE *a, GCGE_INT *lda,
            GCGE_DOUBLE *b, GCGE_INT *ldb, GCGE_DOUBLE *beta, GCGE_DOUBLE *c, GCGE_INT *ldc)
{
    LAPACK_CALL(dsymm_, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}

GCGE_DOUBLE d_dot(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT n)
{
    int inc = 1;
    return lapack_dot(&n, x, &inc, y, &inc);
}

GCGE_DOUBLE d_nrm2(GCGE_DOUBLE *x, GCGE_INT n)
{
    int inc = 1;
    return lapack_nrm2(&n, x, &inc);
}

// y = a x + b y as GCGE_Default_ArrayAXPBY (gcge_ops.c:391-405)
void d_axpby(GCGE_DOUBLE a, GCGE_DOUBLE *x, GCGE_DOUBLE b, GCGE_DOUBLE *y, GCGE_INT n)
{
    int inc = 1;
    if (b != 1.0) LAPACK_CALL(dscal_, &n, &b, y, &inc);
    LAPACK_CALL(daxpy_, &n, &a, x, &inc, y, &inc);
}

void d_copy(GCGE_DOUBLE *x, GCGE_DOUBLE *y, GCGE_INT n)
{
    int inc = 1;
    LAPACK_CALL(dcopy_, &n, x, &inc, y, &inc);
}

void d_scale(GCGE_DOUBLE a, GCGE_DOUBLE *x, GCGE_INT n)
{
    int inc = 1;
    LAPACK_CALL(dscal_, &n, &a, x, &inc);
}

}  // namespace

extern "C" void GCGE_SetDenseOps(GCGE_OPS *ops)
{
    ops->DenseMatEigenSolver = d_eig;
    ops->DenseMatDotDenseMat = d_gemm;
    ops->DenseSymMatDotDenseMat = d_symm;
    ops->ArrayDotArray = d_dot;
    ops->ArrayNorm = d_nrm2;
    ops->ArrayAXPBY = d_axpby;
    ops->ArrayCopy = d_copy;
    ops->ArrayScale = d_scale;
}

// Upper Cholesky factor S = R^T R of the n x n symmetric positive definite
// matrix S (column-major, leading dimension lds; upper triangle read), then
// R^{-1} in place of the upper triangle.  Returns LAPACK's info: 0 on
// success, k > 0 when the leading minor of order k is not positive
// definite (column k-1 depends on the columns before it), -1 when the
// library lacks the routines.
int gcge_chol_inverse(int n, double *S, int lds)
{
    Lapack &L = lapack();
    if (!L.ok || !L.dpotrf_ || !L.dtrtri_) return -1;
    int info = 0;
    L.dpotrf_("U", &n, S, &lds, &info);
    if (info != 0) return info;
    L.dtrtri_("U", "N", &n, S, &lds, &info);
    return info;
}

// All eigenpairs of the n x n symmetric matrix S (column-major, leading
// dimension lds, upper triangle read; S is overwritten): eigenvalues in
// ascending order in w[n], orthonormal eigenvectors in Z (n x n, ldz = n).
// Returns LAPACK's info (0 on success), -1 without LAPACK.
int gcge_sym_eig(int n, double *S, int lds, double *w, double *Z)
{
    Lapack &L = lapack();
    if (!L.ok) return -1;
    double vl = 0.0, vu = 0.0, abstol = 0.0;
    int il = 1, iu = n, m = 0, info = 0, lwork = 8 * n, ldz = n;
    double *work = (double *)malloc(sizeof(double) * (size_t)lwork);
    int *iwork = (int *)malloc(sizeof(int) * 5 * (size_t)n);
    int *ifail = (int *)malloc(sizeof(int) * (size_t)n);
    L.dsyevx_("V", "A", "U", &n, S, &lds, &vl, &vu, &il, &iu, &abstol, &m, w, Z, &ldz, work, &lwork, iwork, ifail,
              &info);
    free(work);
    free(iwork);
    free(ifail);
    return info;
}

// Eigenpairs il..iu (1-based, ascending) of the n x n symmetric matrix S
// (column-major, leading dimension lds, upper triangle read; S is left
// unchanged) from its full decomposition by dsyevd (divide and conquer):
// eigenvalues to w[0 .. iu-il], eigenvectors to the columns of Z (leading
// dimension ldz).  Falls back to dsyevx with range "I" when the library has
// no dsyevd.  Returns LAPACK's info (0 on success), -1 without LAPACK.
int gcge_sym_eig_range(int n, const double *S, int lds, int il, int iu, double *w, double *Z, int ldz)
{
    Lapack &L = lapack();
    if (!L.ok || n <= 0 || il < 1 || iu > n || il > iu) return -1;
    thread_local std::vector<double> a, wa, work;
    thread_local std::vector<int> iwork;
    a.resize((size_t)n * n);
    for (int j = 0; j < n; j++) std::memcpy(a.data() + (size_t)j * n, S + (size_t)j * lds, sizeof(double) * n);
    int nn = n, lda = n, info = 0;
    const int m = iu - il + 1;
    if (!L.dsyevd_) {
        double vl = 0.0, vu = 0.0, abstol = 0.0;
        int mm = 0, lw = 8 * n, ldzz = ldz, ilo = il, ihi = iu;
        work.resize(lw);
        iwork.resize(6 * (size_t)n);
        wa.resize(n);
        L.dsyevx_("V", "I", "U", &nn, a.data(), &lda, &vl, &vu, &ilo, &ihi, &abstol, &mm, wa.data(), Z, &ldzz,
                  work.data(), &lw, iwork.data(), iwork.data() + 5 * (size_t)n, &info);
        std::memcpy(w, wa.data(), sizeof(double) * m);
        return info;
    }
    int lwork = 1 + 6 * n + 2 * n * n, liwork = 3 + 5 * n;
    work.resize(lwork);
    iwork.resize(liwork);
    wa.resize(n);
    L.dsyevd_("V", "U", &nn, a.data(), &lda, wa.data(), work.data(), &lwork, iwork.data(), &liwork, &info);
    if (info != 0) return info;
    for (int j = 0; j < m; j++) {
        w[j] = wa[il - 1 + j];
        std::memcpy(Z + (size_t)j * ldz, a.data() + (size_t)(il - 1 + j) * n, sizeof(double) * n);
    }
    return 0;
}
