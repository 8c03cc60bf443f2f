// hip_common.h -- internal types of the gfx950 GCGE backend.
//
// One process drives one GPU through one HIP stream (default priority).
// Multivectors are single column-major allocations whose leading dimension
// is padded to an odd multiple of 256 elements, so every column starts on a
// 2 KiB (FP64) / 1 KiB (FP32) boundary and any column range can be streamed
// with aligned, coalesced loads.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "gcge_app_hip.h"

namespace gh {

constexpr uint32_t kMagicMV = 0x4d56434fu;   // multivector
constexpr uint32_t kMagicCol = 0x4c4f4356u;  // column view / vector
constexpr uint32_t kMagicMat = 0x54414d48u;  // HIP_MAT

constexpr int kMaxCols = 128;     // per-launch column batch for per-column scalars (2.5 KB of kernel arguments)
constexpr int kMaxChunks = 512;   // split-n partial sums of reductions

struct MV;

// A vector handle (void *) is a view of one column of a multivector.
struct Col {
    uint32_t magic;
    MV *mv;
    int col;
    int owner;  // 1: this vector owns mv (created by VecCreateBy*)
};

struct MV {
    uint32_t magic;
    int prec;         // GCGE_HIP_FP64 / GCGE_HIP_FP32
    int64_t n;        // rows
    int64_t ld;       // leading dimension in elements
    int ncols;
    void *data;       // device memory, ld * ncols elements
    Col *cols;        // ncols column views
    HIP_MAT *mat;     // matrix this multivector was created from (may be NULL)
};

}  // namespace gh

// Rows with more than this many stored entries are multiplied by a whole
// wavefront each (k_spmm_long) instead of one thread (k_spmm): power-law
// graphs carry rows thousands of entries long.
constexpr int kLongRow = 64;

struct HIP_MAT_ {
    uint32_t magic;
    int n;
    int nnz;
    int *rowptr;      // device, n+1
    int *col;         // device, nnz
    int *rowidx;      // device, nnz (row of each stored entry; built on demand)
    double *val64;    // device, nnz
    float *val32;     // device, nnz (built on demand)
    int nlong;        // rows with more than kLongRow stored entries
    int *long_rows;   // device, nlong: their indices, ascending (null when none)
    double *dinv64;   // device, n: 1 / a_ii (1 where a_ii is 0 or absent; built on demand)
    float *dinv32;    // device, n: the same, rounded to FP32
    int prec;         // precision of multivectors created from this matrix
    int max_row_nnz;
    int gather_rows;  // SpMM gathers from a row-major copy of X (scattered column indices)
    double row_abs_max;  // max_i sum_j |a_ij| (Gershgorin bound of ||A||_2)
    // staged-row-block SpMM (k_spmm_halo), chosen at creation when every
    // 256-row block reads at most kHaloCap - 256 rows besides its own
    int halo;                 // 1: SpMM by k_spmm_halo
    int halo_cap;             // LDS slots per column: 256 own rows + the most extras of any block (rounded)
    int *halo_ptr;            // device, nblk+1: each block's extras in halo_col
    int *halo_col;            // device: the other rows a block's entries read, ascending
    unsigned short *lcol;     // device, nnz: each entry's slot (own row r0+i -> i, extra j -> 256+j)
    double halo_bytes;        // device bytes of the three arrays
    // row order (HIP_MatCreateFromCSROrdered): the device matrix is
    // P A P^T; multivectors created from it hold rows in the same order and
    // are permuted at host transfers.  Host arrays, null when unordered.
    int *perm_h;              // n: new position -> the caller's row
    int *iperm_h;             // n: the caller's row -> new position
    int *nzmap_h;             // nnz: the caller's entry index -> stored entry index
};

// largest LDS slot count per column of a k_spmm_halo row block
constexpr int kHaloCap = 4096;

namespace gh {

struct Ctx {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    // device scratch for partial reductions and staging (grown on demand)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    void *rows = nullptr;  // row-major X staging of the SpMM gather path
    size_t rows_bytes = 0;
    // pinned host buffer for results read back by the host
    double *hbuf = nullptr;
    size_t hbuf_bytes = 0;
    double dev_bytes = 0.0;
    int num_cu = 0;
    // host->device staging ring (pinned host side, device side)
    char *stage_h = nullptr;
    char *stage_d = nullptr;
    size_t stage_cap = 0;
    size_t stage_off = 0;
};

Ctx &ctx();
bool ensure_ctx();
// device address of this GPU's no-allocation scratch array (noalloc.hip),
// looked up once per device at context set-up
bool noalloc_setup(int device);
void *scratch(size_t bytes);           // device scratch of at least bytes
void *row_stage(size_t bytes);         // device staging of row-major X copies (SpMM), at least bytes
double *host_buf(size_t count);        // pinned host doubles
// Copy small host data (coefficients) to device memory without blocking:
// the bytes go through a pinned ring whose slots are reused only after the
// next stream synchronisation.  Returns the device address.
void *stage_to_device(const void *host, size_t bytes);
bool sync();                           // stream synchronize + error check

// ---- kernel profiler (opt-in): HIP events around each launch of a kernel
// class on the backend stream, with the algorithmic bytes and flops of the
// launch; resolved at stream synchronisation points.
enum KClass {
    K_SPMM_F64 = 0, K_SPMM_F32, K_SPMM_DOT_F64, K_SPMM_DOT_F32, K_AXPBY, K_DOTS, K_CG_UPDATE_F64, K_CG_UPDATE_F32,
    K_CG_SEARCH_F64, K_CG_SEARCH_F32,
    K_GRAM_MFMA_F64, K_GRAM_MFMA_F32, K_GRAM_THIN, K_LINCOMB_MFMA_F64, K_LINCOMB_MFMA_F32, K_LINCOMB_THIN,
    K_RESIDUAL, K_COPY, K_SWAP, K_OTHER, K_NUM
};
struct ProfScope {
    int kind;
    bool on;
    ProfScope(int kind, double bytes, double flops);
    ~ProfScope();
};
const char *kclass_name(int k);

// error latch
void latch(int code, const char *fmt, ...);
bool failed();
bool check(hipError_t e, const char *what);

inline MV *as_mv(void **h) {
    MV *m = reinterpret_cast<MV *>(h);
    return (m && m->magic == kMagicMV) ? m : nullptr;
}
inline Col *as_col(void *h) {
    Col *c = reinterpret_cast<Col *>(h);
    return (c && c->magic == kMagicCol) ? c : nullptr;
}
inline HIP_MAT *as_mat(void *h) {
    HIP_MAT *m = reinterpret_cast<HIP_MAT *>(h);
    return (m && m->magic == kMagicMat) ? m : nullptr;
}

size_t elem_size(int prec);
void *col_ptr(MV *m, int64_t c);
const float *mat_val32(HIP_MAT *A);
const int *mat_rowidx(HIP_MAT *A);
// 1 / diagonal of A in the given precision (built on first use)
const void *mat_dinv(HIP_MAT *A, int prec);

// ---- launchers (hip_kernels.hip) --------------------------------------
// All launchers validate nothing; callers check ranges first.
// Column j of a range is base + j * ld.

// y_j = a_j x_j + b_j y_j (branch rules of y = a x + b y), j < k <= kMaxCols
void launch_axpby(int prec, int64_t n, const void *x, int64_t ldx, void *y, int64_t ldy, int k,
                  const double *a, const double *b, const int *xidx);
// out_j = <x_j, y_j>, FP64 accumulation; result in host array out
bool dots(int prec, int64_t n, const void *x, int64_t ldx, const void *y, int64_t ldy, int k,
          double *out);
// y_j <- x_j (with precision conversion), j < k
void launch_copy(int xprec, const void *x, int64_t ldx, int yprec, void *y, int64_t ldy, int64_t n,
                 int k);
void launch_swap(int prec, void *x, int64_t ldx, void *y, int64_t ldy, int64_t n, int k);
void launch_zero_rows(int prec, void *x, int64_t ldx, int64_t n, int k, const unsigned char *mask);
// Y_j = A X_j (xidx: optional gather of X columns, k <= kMaxCols when given)
void launch_spmm(int prec, HIP_MAT *A, const void *X, int64_t ldx, void *Y, int64_t ldy, int k,
                 const int *xidx);
// Y_j = A X_j and out_j = <X_j, Y_j>
bool spmm_dot(int prec, HIP_MAT *A, const void *X, int64_t ldx, void *Y, int64_t ldy, int k,
              double *out);
// rnorm/xnorm/bxnorm of A x_j - lambda_j B x_j, x_j = X(:, idx[j])
bool residual_norms(int prec, HIP_MAT *A, HIP_MAT *B, const void *X, int64_t ldx, const int *idx,
                    int k, const double *lambda, double *rnorm, double *xnorm, double *bxnorm);
// x_j += alpha_j p_j ; r_j -= alpha_j w_j ; rr_j = <r_j, r_j>
bool cg_update(int prec, int64_t n, void *x, int64_t ldx, const void *p, int64_t ldp, const void *w,
               int64_t ldw, void *r, int64_t ldr, int k, const double *alpha, double *rr);
// r_j -= alpha_j w_j ; rr_j = <r_j, r_j>
bool cg_resid(int prec, int64_t n, const void *w, int64_t ldw, void *r, int64_t ldr, int k, const double *alpha,
              double *rr);
// r_j -= alpha_j w_j ; rz_j = <r_j, dinv .* r_j> ; rr_j = <r_j, r_j>
bool pcg_resid(int prec, int64_t n, const void *w, int64_t ldw, void *r, int64_t ldr, const void *dinv, int k,
               const double *alpha, double *rz, double *rr);
// dinv64[i] = 1 / a_ii (1 where a_ii is 0 or absent), dinv32 its FP32 rounding
void launch_diag_inv(HIP_MAT *A);
// x_j += alpha_j p_j (j < kx), then p_j = r_j + beta_j p_j (j < kp); with
// dinv (not null), p_j = dinv .* r_j + beta_j p_j
void launch_cg_search(int prec, int64_t n, void *x, int64_t ldx, void *p, int64_t ldp, const void *r, int64_t ldr,
                      int kx, int kp, const double *alpha, const double *beta, const void *dinv = nullptr);
// G (p x q, host, column-major ldg) = V^T W
bool gram(int prec, int64_t n, const void *V, int64_t ldv, int p, const void *W, int64_t ldw, int q,
          double *G, int ldg, bool upper_only);
// Y = beta Y + alpha X C  (C host, p x q, column-major ldc)
bool lincomb(int prec, int64_t n, const void *X, int64_t ldx, int p, void *Y, int64_t ldy, int q,
             const double *C, int ldc, double alpha, double beta);
// per-nonzero sampled outer product g_e = sum_i w_i x_i[row_e] x_i[col_e]
bool sampled_outer(int prec, HIP_MAT *A, const void *X, int64_t ldx, int k, const double *w,
                   double *g_dev);
// grad[p] = scale * sum_t jval[t] g[jnz[t]] over t in [jrowptr[p], jrowptr[p+1]),
// added to grad[p] when accumulate
void launch_csr_gather_sum(int nparam, const int *jrowptr, const int *jnz, const double *jval,
                           const double *g, double *grad, double scale, bool accumulate);

}  // namespace gh
