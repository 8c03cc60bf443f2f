// hip_backend.cpp -- "app/hip": device matrices, multivectors and the
// GCGE_OPS slots of the gfx950 backend.
//
// The slot set mirrors app/csr (app/csr/gcge_app_csr.c:26-90) plus every
// multivector slot, which app/csr leaves to the per-column defaults of
// src/gcge_ops.c.  Host code validates every column range against the
// multivector it addresses before launching, and latches the first error
// (the slots return void).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstddef>
#include <cstring>
#include <vector>

#include "gcge_app_hip.h"
#include "hip_common.h"

namespace gh {
void launch_rowidx(HIP_MAT *A);
void launch_to_f32(int64_t n, const double *a, float *b);

// ------------------------------------------------------------ error latch
static int g_err = 0;
static char g_msg[512] = {0};

void latch(int code, const char *fmt, ...)
{
    if (g_err) return;
    g_err = code ? code : 1;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    if (getenv("GCGE_HIP_VERBOSE")) fprintf(stderr, "gcge-hip error: %s\n", g_msg);
}

bool failed() { return g_err != 0; }

bool check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        latch((int)e, "%s: %s", what, hipGetErrorString(e));
        return false;
    }
    return true;
}

// ------------------------------------------------------------ context
static Ctx g_ctx;

Ctx &ctx() { return g_ctx; }

bool ensure_ctx()
{
    if (g_ctx.init) return true;
    int ndev = 0;
    if (!check(hipGetDeviceCount(&ndev), "hipGetDeviceCount")) return false;
    if (ndev < 1) {
        latch(100, "no HIP device");
        return false;
    }
    if (!check(hipGetDevice(&g_ctx.device), "hipGetDevice")) return false;
    if (!check(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking), "hipStreamCreate")) return false;
    hipDeviceProp_t prop;
    if (check(hipGetDeviceProperties(&prop, g_ctx.device), "hipGetDeviceProperties"))
        g_ctx.num_cu = prop.multiProcessorCount;
    if (!noalloc_setup(g_ctx.device)) return false;
    g_ctx.init = true;
    return true;
}

void *scratch(size_t bytes)
{
    if (!ensure_ctx()) return nullptr;
    if (bytes <= g_ctx.scratch_bytes) return g_ctx.scratch;
    if (g_ctx.scratch) {
        // kernels queued on the stream may still read the old buffer
        if (!sync()) return nullptr;
        (void)hipFree(g_ctx.scratch);
        g_ctx.dev_bytes -= (double)g_ctx.scratch_bytes;
        g_ctx.scratch = nullptr;
        g_ctx.scratch_bytes = 0;
    }
    size_t want = std::max(bytes, (size_t)(8u << 20));
    want = (want + 4095) & ~(size_t)4095;
    if (!check(hipMalloc(&g_ctx.scratch, want), "hipMalloc scratch")) return nullptr;
    g_ctx.scratch_bytes = want;
    g_ctx.dev_bytes += (double)want;
    return g_ctx.scratch;
}

void *row_stage(size_t bytes)
{
    if (!ensure_ctx()) return nullptr;
    if (bytes <= g_ctx.rows_bytes) return g_ctx.rows;
    if (g_ctx.rows) {
        if (!sync()) return nullptr;  // kernels queued on the stream may still read it
        (void)hipFree(g_ctx.rows);
        g_ctx.dev_bytes -= (double)g_ctx.rows_bytes;
        g_ctx.rows = nullptr;
        g_ctx.rows_bytes = 0;
    }
    const size_t want = (bytes + ((size_t)1 << 20) - 1) & ~(((size_t)1 << 20) - 1);
    if (!check(hipMalloc(&g_ctx.rows, want), "hipMalloc row staging")) return nullptr;
    g_ctx.rows_bytes = want;
    g_ctx.dev_bytes += (double)want;
    return g_ctx.rows;
}

double *host_buf(size_t count)
{
    if (!ensure_ctx()) return nullptr;
    size_t bytes = count * sizeof(double);
    if (bytes <= g_ctx.hbuf_bytes) return g_ctx.hbuf;
    if (g_ctx.hbuf) {
        if (!sync()) return nullptr;
        (void)hipHostFree(g_ctx.hbuf);
        g_ctx.hbuf = nullptr;
    }
    size_t want = std::max(bytes, (size_t)(1u << 20));
    if (!check(hipHostMalloc((void **)&g_ctx.hbuf, want, hipHostMallocDefault), "hipHostMalloc")) {
        g_ctx.hbuf_bytes = 0;
        return nullptr;
    }
    g_ctx.hbuf_bytes = want;
    return g_ctx.hbuf;
}

// ------------------------------------------------------------ profiler
struct ProfEntry {
    int kind;
    hipEvent_t a, b;
    double bytes, flops;
};
struct ProfStat {
    long count = 0;
    double ms = 0.0, bytes = 0.0, flops = 0.0;
};
static bool g_prof = false;
static std::vector<ProfEntry> g_pending;
static std::vector<hipEvent_t> g_event_pool;
static ProfStat g_stats[K_NUM];

static hipEvent_t ev_get()
{
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static void prof_resolve()
{
    for (auto &p : g_pending) {
        float ms = 0.f;
        if (p.a && p.b && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ProfStat &st = g_stats[p.kind];
            st.count++;
            st.ms += ms;
            st.bytes += p.bytes;
            st.flops += p.flops;
        }
        if (p.a) g_event_pool.push_back(p.a);
        if (p.b) g_event_pool.push_back(p.b);
    }
    g_pending.clear();
}

const char *kclass_name(int k)
{
    // one class per kernel family and storage precision (k_spmm<T, DOT>,
    // k_cg_update<T>, k_gram_mfma<T, ...>, k_lincomb_mfma<T, ...>, ...)
    static const char *names[K_NUM] = {"spmm_f64", "spmm_f32", "spmm_dot_f64", "spmm_dot_f32", "axpby", "dots",
                                       "cg_update_f64", "cg_update_f32", "cg_search_f64", "cg_search_f32",
                                       "gram_mfma_f64", "gram_mfma_f32",
                                       "gram_thin", "lincomb_mfma_f64", "lincomb_mfma_f32", "lincomb_thin",
                                       "residual", "copy", "swap", "other"};
    return (k >= 0 && k < K_NUM) ? names[k] : "?";
}

// scopes currently open: their end events are not recorded yet, so a
// synchronisation inside one must leave the pending list alone
static int g_prof_open = 0;

ProfScope::ProfScope(int k, double bytes, double flops) : kind(k), on(g_prof && g_ctx.init)
{
    if (!on) return;
    g_prof_open++;
    ProfEntry e;
    e.kind = k;
    e.a = ev_get();
    e.b = ev_get();
    e.bytes = bytes;
    e.flops = flops;
    if (e.a) (void)hipEventRecord(e.a, g_ctx.stream);
    g_pending.push_back(e);
}

ProfScope::~ProfScope()
{
    if (!on) return;
    g_prof_open--;
    if (g_pending.empty()) return;
    ProfEntry &e = g_pending.back();
    if (e.b) (void)hipEventRecord(e.b, g_ctx.stream);
    if (g_pending.size() > 4096) sync();
}

bool sync()
{
    if (!g_ctx.init) return true;
    bool ok = check(hipStreamSynchronize(g_ctx.stream), "hipStreamSynchronize");
    if (!g_pending.empty() && g_prof_open == 0) prof_resolve();
    g_ctx.stage_off = 0;  // every staged copy has completed
    return ok;
}

void *stage_to_device(const void *host, size_t bytes)
{
    if (!ensure_ctx()) return nullptr;
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (need > g_ctx.stage_cap) {
        if (!sync()) return nullptr;
        if (g_ctx.stage_h) (void)hipHostFree(g_ctx.stage_h);
        if (g_ctx.stage_d) {
            (void)hipFree(g_ctx.stage_d);
            g_ctx.dev_bytes -= (double)g_ctx.stage_cap;
        }
        g_ctx.stage_h = nullptr;
        g_ctx.stage_d = nullptr;
        const size_t cap = std::max(need, (size_t)(8u << 20));
        if (!check(hipHostMalloc((void **)&g_ctx.stage_h, cap, hipHostMallocDefault), "hipHostMalloc stage") ||
            !check(hipMalloc((void **)&g_ctx.stage_d, cap), "hipMalloc stage")) {
            g_ctx.stage_cap = 0;
            return nullptr;
        }
        g_ctx.stage_cap = cap;
        g_ctx.dev_bytes += (double)cap;
        g_ctx.stage_off = 0;
    }
    if (g_ctx.stage_off + need > g_ctx.stage_cap && !sync()) return nullptr;
    char *h = g_ctx.stage_h + g_ctx.stage_off;
    char *d = g_ctx.stage_d + g_ctx.stage_off;
    std::memcpy(h, host, bytes);
    if (!check(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, g_ctx.stream), "stage H2D")) return nullptr;
    g_ctx.stage_off += need;
    return d;
}

size_t elem_size(int prec) { return prec == GCGE_HIP_FP32 ? sizeof(float) : sizeof(double); }

void *col_ptr(MV *m, int64_t c) { return (char *)m->data + (size_t)c * (size_t)m->ld * elem_size(m->prec); }

const float *mat_val32(HIP_MAT *A)
{
    if (A->val32) return A->val32;
    if (!check(hipMalloc((void **)&A->val32, sizeof(float) * std::max(1, A->nnz)), "hipMalloc val32"))
        return nullptr;
    g_ctx.dev_bytes += sizeof(float) * (double)A->nnz;
    launch_to_f32(A->nnz, A->val64, A->val32);
    return failed() ? nullptr : A->val32;
}

const void *mat_dinv(HIP_MAT *A, int prec)
{
    if (!A->dinv64) {
        const size_t n = (size_t)std::max(1, A->n);
        if (!check(hipMalloc((void **)&A->dinv64, sizeof(double) * n), "hipMalloc dinv64")) return nullptr;
        if (!check(hipMalloc((void **)&A->dinv32, sizeof(float) * n), "hipMalloc dinv32")) return nullptr;
        g_ctx.dev_bytes += 12.0 * (double)n;
        launch_diag_inv(A);
        if (failed()) return nullptr;
    }
    return prec == GCGE_HIP_FP64 ? (const void *)A->dinv64 : (const void *)A->dinv32;
}

const int *mat_rowidx(HIP_MAT *A)
{
    if (A->rowidx) return A->rowidx;
    if (!check(hipMalloc((void **)&A->rowidx, sizeof(int) * std::max(1, A->nnz)), "hipMalloc rowidx"))
        return nullptr;
    g_ctx.dev_bytes += sizeof(int) * (double)A->nnz;
    launch_rowidx(A);
    return failed() ? nullptr : A->rowidx;
}

// ------------------------------------------------------------ multivectors
// An odd multiple of 256 elements: columns start on 2 KiB (FP64) / 1 KiB
// (FP32) boundaries, and their stride is no multiple of 4 KiB, so the same
// row of neighbouring columns falls in different cache sets and memory
// channels (measured: the staged SpMM 2 % faster than with ld = n for
// n = 4,096,000, a multiple of 2^18 bytes per FP64 column).
static int64_t pad_ld(int64_t n) { return (((n + 255) / 256) | 1) * 256; }

static MV *mv_new(int64_t n, int ncols, int prec, HIP_MAT *mat)
{
    if (!ensure_ctx()) return nullptr;
    if (n <= 0 || ncols < 0) {
        latch(101, "multivector: bad shape n=%lld ncols=%d", (long long)n, ncols);
        return nullptr;
    }
    MV *m = new MV;
    m->magic = kMagicMV;
    m->prec = prec;
    m->n = n;
    m->ld = pad_ld(n);
    m->ncols = ncols;
    m->mat = mat;
    m->data = nullptr;
    size_t bytes = (size_t)m->ld * (size_t)std::max(ncols, 1) * elem_size(prec);
    if (!check(hipMalloc(&m->data, bytes), "hipMalloc multivector")) {
        delete m;
        return nullptr;
    }
    g_ctx.dev_bytes += (double)bytes;
    // vectors start at zero, as the calloc of CSR_VecCreateByMat does
    check(hipMemsetAsync(m->data, 0, bytes, g_ctx.stream), "hipMemset multivector");
    m->cols = new Col[std::max(ncols, 1)];
    for (int j = 0; j < std::max(ncols, 1); j++) {
        m->cols[j].magic = kMagicCol;
        m->cols[j].mv = m;
        m->cols[j].col = j;
        m->cols[j].owner = 0;
    }
    return m;
}

static void mv_free(MV *m)
{
    if (!m || m->magic != kMagicMV) return;
    sync();
    if (m->data) {
        (void)hipFree(m->data);
        g_ctx.dev_bytes -= (double)((size_t)m->ld * (size_t)std::max(m->ncols, 1) * elem_size(m->prec));
    }
    for (int j = 0; j < std::max(m->ncols, 1); j++) m->cols[j].magic = 0;
    delete[] m->cols;
    m->magic = 0;
    delete m;
}

static bool range_ok(MV *m, int s, int e, const char *what)
{
    if (!m) {
        latch(102, "%s: not a device multivector", what);
        return false;
    }
    if (s < 0 || e < s || e > m->ncols) {
        latch(103, "%s: column range [%d,%d) outside [0,%d)", what, s, e, m->ncols);
        return false;
    }
    return true;
}

static bool same_rows(MV *a, MV *b, const char *what)
{
    if (a->n != b->n) {
        latch(104, "%s: row counts differ (%lld vs %lld)", what, (long long)a->n, (long long)b->n);
        return false;
    }
    return true;
}

static bool same_prec(MV *a, MV *b, const char *what)
{
    if (a->prec != b->prec) {
        latch(105, "%s: precisions differ", what);
        return false;
    }
    return true;
}

static bool mat_ok(HIP_MAT *A, MV *x, const char *what)
{
    if (!A || A->magic != kMagicMat) {
        latch(106, "%s: not a HIP_MAT", what);
        return false;
    }
    if (A->n != x->n) {
        latch(107, "%s: matrix has %d rows, vectors %lld", what, A->n, (long long)x->n);
        return false;
    }
    return true;
}

}  // namespace gh

using namespace gh;

// ====================================================================== matrices
static int pick_gather_rows(int n, const int *rowptr, const int *col);
static bool build_halo(HIP_MAT *A, const int *rowptr, const int *col);

extern "C" HIP_MAT *HIP_MatCreateFromCSR(GCGE_INT nrows, GCGE_INT ncols, GCGE_INT nnz, const GCGE_INT *rowptr,
                                         const GCGE_INT *col, const GCGE_DOUBLE *val)
{
    if (!ensure_ctx()) return nullptr;
    if (nrows <= 0 || ncols != nrows || nnz < 0 || !rowptr || (nnz > 0 && (!col || !val))) {
        latch(110, "HIP_MatCreateFromCSR: bad arguments (n=%d, m=%d, nnz=%d)", nrows, ncols, nnz);
        return nullptr;
    }
    if (rowptr[0] != 0 || rowptr[nrows] != nnz) {
        latch(111, "HIP_MatCreateFromCSR: rowptr[0]=%d rowptr[n]=%d nnz=%d", rowptr[0], rowptr[nrows], nnz);
        return nullptr;
    }
    int maxrow = 0;
    double rowabs = 0.0;
    for (int i = 0; i < nrows; i++) {
        int a = rowptr[i], b = rowptr[i + 1];
        if (b < a) {
            latch(112, "HIP_MatCreateFromCSR: rowptr decreases at row %d", i);
            return nullptr;
        }
        maxrow = std::max(maxrow, b - a);
        double s = 0.0;
        for (int e = a; e < b; e++) {
            if (col[e] < 0 || col[e] >= ncols) {
                latch(113, "HIP_MatCreateFromCSR: column index %d out of range in row %d", col[e], i);
                return nullptr;
            }
            s += std::fabs(val[e]);
        }
        rowabs = std::max(rowabs, s);
    }
    HIP_MAT *A = new HIP_MAT_;
    std::memset(A, 0, sizeof(*A));
    A->magic = kMagicMat;
    A->n = nrows;
    A->nnz = nnz;
    A->prec = GCGE_HIP_FP64;
    A->max_row_nnz = maxrow;
    A->row_abs_max = rowabs;
    std::vector<int> longs;
    for (int i = 0; i < nrows; i++)
        if (rowptr[i + 1] - rowptr[i] > kLongRow) longs.push_back(i);
    A->nlong = (int)longs.size();
    A->gather_rows = pick_gather_rows(nrows, rowptr, col);
    bool ok = check(hipMalloc((void **)&A->rowptr, sizeof(int) * (nrows + 1)), "hipMalloc rowptr") &&
              check(hipMalloc((void **)&A->col, sizeof(int) * std::max(1, nnz)), "hipMalloc col") &&
              check(hipMalloc((void **)&A->val64, sizeof(double) * std::max(1, nnz)), "hipMalloc val");
    if (ok) {
        ok = check(hipMemcpy(A->rowptr, rowptr, sizeof(int) * (nrows + 1), hipMemcpyHostToDevice), "H2D rowptr") &&
             (nnz == 0 || check(hipMemcpy(A->col, col, sizeof(int) * nnz, hipMemcpyHostToDevice), "H2D col")) &&
             (nnz == 0 || check(hipMemcpy(A->val64, val, sizeof(double) * nnz, hipMemcpyHostToDevice), "H2D val"));
    }
    if (ok && A->nlong > 0)
        ok = check(hipMalloc((void **)&A->long_rows, sizeof(int) * A->nlong), "hipMalloc long_rows") &&
             check(hipMemcpy(A->long_rows, longs.data(), sizeof(int) * A->nlong, hipMemcpyHostToDevice),
                   "H2D long_rows");
    if (ok && !A->gather_rows) ok = build_halo(A, rowptr, col);
    if (!ok) {
        HIP_MatDestroy(&A);
        return nullptr;
    }
    g_ctx.dev_bytes += sizeof(int) * (double)(nrows + 1) + 12.0 * nnz;
    return A;
}

// Staged row blocks of k_spmm_halo: for each 256-row block, the rows its
// entries read besides its own (ascending), and every entry's LDS slot.
// Chosen for the column-gather layout when the blocks are compact: on
// average a block reads at most as many other rows as it owns, and no block
// more than kHaloCap - 256 (a grid in tile order reads its surface, about
// 0.9 times its rows; the natural order of a 3-D grid reads whole planes,
// 2.4 times, where the direct gather of k_spmm is as fast).
// GCGE_HIP_SPMM_GATHER=halo takes it whenever it fits, =columns or =rows
// never.  Rows longer than kLongRow (k_spmm_long) are left out.
static bool build_halo(HIP_MAT *A, const int *rowptr, const int *col)
{
    const char *e = getenv("GCGE_HIP_SPMM_GATHER");
    const bool forced = e && !strcmp(e, "halo");
    if (e && !forced) return true;
    const int n = A->n, R = 256, nblk = (n + R - 1) / R;
    std::vector<int> hptr((size_t)nblk + 1, 0), hcol, ext;
    std::vector<unsigned short> lcol((size_t)std::max(1, A->nnz), 0);
    int most = 0;
    for (int b = 0; b < nblk; b++) {
        const int r0 = b * R, r1 = std::min(n, r0 + R);
        ext.clear();
        for (int r = r0; r < r1; r++) {
            if (rowptr[r + 1] - rowptr[r] > kLongRow) continue;
            for (int q = rowptr[r]; q < rowptr[r + 1]; q++)
                if (col[q] < r0 || col[q] >= r1) ext.push_back(col[q]);
        }
        std::sort(ext.begin(), ext.end());
        ext.erase(std::unique(ext.begin(), ext.end()), ext.end());
        if ((int)ext.size() > kHaloCap - R) return true;  // too wide for LDS: keep k_spmm
        most = std::max(most, (int)ext.size());
        for (int r = r0; r < r1; r++) {
            if (rowptr[r + 1] - rowptr[r] > kLongRow) continue;
            for (int q = rowptr[r]; q < rowptr[r + 1]; q++) {
                const int c = col[q];
                lcol[q] = (unsigned short)((c >= r0 && c < r1)
                                               ? c - r0
                                               : R + (int)(std::lower_bound(ext.begin(), ext.end(), c) - ext.begin()));
            }
        }
        hcol.insert(hcol.end(), ext.begin(), ext.end());
        hptr[b + 1] = (int)hcol.size();
        if (!forced && (int64_t)hcol.size() > (int64_t)n) return true;  // not compact: keep k_spmm
    }
    const size_t nh = std::max<size_t>(1, hcol.size());
    bool ok = check(hipMalloc((void **)&A->halo_ptr, sizeof(int) * hptr.size()), "hipMalloc halo_ptr") &&
              check(hipMalloc((void **)&A->halo_col, sizeof(int) * nh), "hipMalloc halo_col") &&
              check(hipMalloc((void **)&A->lcol, sizeof(unsigned short) * lcol.size()), "hipMalloc lcol");
    if (!ok) return false;
    ok = check(hipMemcpy(A->halo_ptr, hptr.data(), sizeof(int) * hptr.size(), hipMemcpyHostToDevice), "H2D halo") &&
         (hcol.empty() ||
          check(hipMemcpy(A->halo_col, hcol.data(), sizeof(int) * hcol.size(), hipMemcpyHostToDevice), "H2D halo")) &&
         check(hipMemcpy(A->lcol, lcol.data(), sizeof(unsigned short) * lcol.size(), hipMemcpyHostToDevice),
               "H2D lcol");
    if (!ok) return false;
    A->halo = 1;
    A->halo_cap = ((R + most + 31) / 32) * 32;
    A->halo_bytes = 4.0 * (double)(hptr.size() + nh) + 2.0 * (double)lcol.size();
    g_ctx.dev_bytes += A->halo_bytes;
    return true;
}

// Gather layout of the SpMM.  The row-major path is used when
// - the rows average more than 8 stored entries: the column-major kernel
//   then cannot stage a 256-row workgroup's entries in LDS and each thread
//   reads its own row's indices and values, 64 cache lines per load (the
//   27-point mass matrix: 3-7x slower SpMM, 15-20x slower residual check);
// - or the column indices scatter: a wavefront of the column-major kernel
//   reads, per stored entry and column, the cache lines its 64 rows' column
//   indices fall in; a stencil's rows share a few lines (about 0.03 distinct
//   32-column segments per entry over a 64-row block), a random graph's do
//   not (0.4 for a Chung-Lu graph); the threshold is 0.15.
// GCGE_HIP_SPMM_GATHER=rows|columns|halo overrides the choice for matrices
// created afterwards (columns: the direct-gather k_spmm; halo: the staged
// row blocks of k_spmm_halo where they fit).
static int pick_gather_rows(int n, const int *rowptr, const int *col)
{
    const char *e = getenv("GCGE_HIP_SPMM_GATHER");
    if (e && !strcmp(e, "rows")) return 1;
    if (e && (!strcmp(e, "columns") || !strcmp(e, "halo"))) return 0;
    if ((double)rowptr[n] > 8.0 * n) return 1;
    const int nb = n / 64;
    if (nb == 0) return 0;
    const int step = std::max(1, nb / 2048);
    std::vector<int> seg;
    double lines = 0.0, ents = 0.0;
    for (int b = 0; b < nb; b += step) {
        seg.assign(col + rowptr[b * 64], col + rowptr[b * 64 + 64]);
        for (int &x : seg) x >>= 5;
        std::sort(seg.begin(), seg.end());
        lines += (double)(std::unique(seg.begin(), seg.end()) - seg.begin());
        ents += (double)seg.size();
    }
    return ents > 0.0 && lines > 0.15 * ents ? 1 : 0;
}

// ---- matrix files: the text format of CSR_ReadMatFile (app/csr/gcge_app_csr.c:185-236)
// and a binary form of the same arrays ("GCGECSR1", int32 nrow, ncol, nnz,
// int32 0, rowptr[nrow+1] int32, col[nnz] int32, val[nnz] f64, little endian)
static const char kCsrMagic[8] = {'G', 'C', 'G', 'E', 'C', 'S', 'R', '1'};

namespace {
struct CsrHost {
    int nrow = 0, ncol = 0, nnz = 0;
    std::vector<int> ia, ja;
    std::vector<double> aa;
};

bool read_all(FILE *f, std::vector<char> &buf)
{
    if (fseek(f, 0, SEEK_END) != 0) return false;
    const long sz = ftell(f);
    if (sz < 0 || fseek(f, 0, SEEK_SET) != 0) return false;
    buf.resize((size_t)sz + 1);
    if (sz > 0 && fread(buf.data(), 1, (size_t)sz, f) != (size_t)sz) return false;
    buf[(size_t)sz] = '\0';
    return true;
}

// whitespace-separated numbers, parsed in place (strtol / strtod: the values
// fscanf's %d / %lf give)
bool parse_text(const std::vector<char> &buf, CsrHost &m, const char *filename)
{
    const char *p = buf.data();
    char *q = nullptr;
    long v[3];
    for (int i = 0; i < 3; i++) {
        v[i] = strtol(p, &q, 10);
        if (q == p) {
            latch(121, "HIP_ReadMatFile: bad header in %s", filename);
            return false;
        }
        p = q;
    }
    if (v[0] <= 0 || v[1] <= 0 || v[2] < 0 || v[0] > INT32_MAX - 1 || v[2] > INT32_MAX) {
        latch(121, "HIP_ReadMatFile: bad header in %s", filename);
        return false;
    }
    m.nrow = (int)v[0];
    m.ncol = (int)v[1];
    m.nnz = (int)v[2];
    m.ia.resize((size_t)m.nrow + 1);
    m.ja.resize((size_t)m.nnz);
    m.aa.resize((size_t)m.nnz);
    for (size_t i = 0; i < m.ia.size(); i++, p = q) {
        m.ia[i] = (int)strtol(p, &q, 10);
        if (q == p) goto truncated;
    }
    for (size_t i = 0; i < m.ja.size(); i++, p = q) {
        m.ja[i] = (int)strtol(p, &q, 10);
        if (q == p) goto truncated;
    }
    for (size_t i = 0; i < m.aa.size(); i++, p = q) {
        m.aa[i] = strtod(p, &q);
        if (q == p) goto truncated;
    }
    return true;
truncated:
    latch(122, "HIP_ReadMatFile: truncated file %s", filename);
    return false;
}

bool parse_binary(const std::vector<char> &buf, CsrHost &m, const char *filename)
{
    const size_t hdr = 8 + 4 * sizeof(int);
    if (buf.size() - 1 < hdr) {
        latch(122, "HIP_ReadMatFile: truncated file %s", filename);
        return false;
    }
    int h[4];
    std::memcpy(h, buf.data() + 8, sizeof(h));
    if (h[0] <= 0 || h[1] <= 0 || h[2] < 0) {
        latch(121, "HIP_ReadMatFile: bad header in %s", filename);
        return false;
    }
    m.nrow = h[0];
    m.ncol = h[1];
    m.nnz = h[2];
    const size_t need = hdr + sizeof(int) * ((size_t)m.nrow + 1) + sizeof(int) * (size_t)m.nnz +
                        sizeof(double) * (size_t)m.nnz;
    if (buf.size() - 1 < need) {
        latch(122, "HIP_ReadMatFile: truncated file %s", filename);
        return false;
    }
    const char *p = buf.data() + hdr;
    m.ia.resize((size_t)m.nrow + 1);
    m.ja.resize((size_t)m.nnz);
    m.aa.resize((size_t)m.nnz);
    std::memcpy(m.ia.data(), p, sizeof(int) * m.ia.size());
    p += sizeof(int) * m.ia.size();
    std::memcpy(m.ja.data(), p, sizeof(int) * m.ja.size());
    p += sizeof(int) * m.ja.size();
    std::memcpy(m.aa.data(), p, sizeof(double) * m.aa.size());
    return true;
}
}  // namespace

extern "C" HIP_MAT *HIP_ReadMatFile(const char *filename)
{
    FILE *f = filename ? fopen(filename, "rb") : nullptr;
    if (!f) {
        latch(120, "HIP_ReadMatFile: cannot open %s", filename ? filename : "(null)");
        return nullptr;
    }
    std::vector<char> buf;
    const bool ok = read_all(f, buf);
    fclose(f);
    if (!ok) {
        latch(120, "HIP_ReadMatFile: cannot read %s", filename);
        return nullptr;
    }
    CsrHost m;
    const bool binary = buf.size() - 1 >= 8 && std::memcmp(buf.data(), kCsrMagic, 8) == 0;
    if (!(binary ? parse_binary(buf, m, filename) : parse_text(buf, m, filename))) return nullptr;
    buf.clear();
    buf.shrink_to_fit();
    return HIP_MatCreateFromCSR(m.nrow, m.ncol, m.nnz, m.ia.data(), m.ja.data(), m.aa.data());
}

extern "C" GCGE_INT HIP_WriteMatFile(const char *filename, GCGE_INT nrows, GCGE_INT ncols, GCGE_INT nnz,
                                     const GCGE_INT *rowptr, const GCGE_INT *col, const GCGE_DOUBLE *val,
                                     GCGE_INT binary)
{
    if (!filename || nrows <= 0 || ncols <= 0 || nnz < 0 || !rowptr || (nnz > 0 && (!col || !val))) {
        latch(123, "HIP_WriteMatFile: bad arguments");
        return 123;
    }
    FILE *f = fopen(filename, binary ? "wb" : "w");
    if (!f) {
        latch(124, "HIP_WriteMatFile: cannot create %s", filename);
        return 124;
    }
    bool ok = true;
    if (binary) {
        const int h[4] = {nrows, ncols, nnz, 0};
        ok = fwrite(kCsrMagic, 1, 8, f) == 8 && fwrite(h, sizeof(int), 4, f) == 4 &&
             fwrite(rowptr, sizeof(int), (size_t)nrows + 1, f) == (size_t)nrows + 1 &&
             fwrite(col, sizeof(int), (size_t)nnz, f) == (size_t)nnz &&
             fwrite(val, sizeof(double), (size_t)nnz, f) == (size_t)nnz;
    } else {
        // one number per line, values with 17 significant digits (they read back exactly)
        ok = fprintf(f, "%d\n%d\n%d\n", nrows, ncols, nnz) > 0;
        for (int i = 0; ok && i <= nrows; i++) ok = fprintf(f, "%d\n", rowptr[i]) > 0;
        for (int i = 0; ok && i < nnz; i++) ok = fprintf(f, "%d\n", col[i]) > 0;
        for (int i = 0; ok && i < nnz; i++) ok = fprintf(f, "%.17g\n", val[i]) > 0;
    }
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
        latch(124, "HIP_WriteMatFile: write to %s failed", filename);
        return 124;
    }
    return 0;
}

extern "C" HIP_MAT *HIP_MatCreateFromCSROrdered(GCGE_INT nrows, GCGE_INT ncols, GCGE_INT nnz, const GCGE_INT *rowptr,
                                                const GCGE_INT *col, const GCGE_DOUBLE *val, const GCGE_INT *perm)
{
    if (nrows <= 0 || ncols != nrows || nnz < 0 || !rowptr || (nnz > 0 && (!col || !val)) ||
        rowptr[nrows] != nnz) {
        latch(110, "HIP_MatCreateFromCSROrdered: bad arguments (n=%d, m=%d, nnz=%d)", nrows, ncols, nnz);
        return nullptr;
    }
    for (int i = 0; i < nrows; i++)
        if (rowptr[i + 1] < rowptr[i] || rowptr[i] < 0) {
            latch(112, "HIP_MatCreateFromCSROrdered: rowptr decreases at row %d", i);
            return nullptr;
        }
    for (int e = 0; e < nnz; e++)
        if (col[e] < 0 || col[e] >= ncols) {
            latch(113, "HIP_MatCreateFromCSROrdered: column index %d out of range", col[e]);
            return nullptr;
        }
    std::vector<int> p((size_t)nrows), ip((size_t)nrows, -1);
    if (perm) {
        std::copy(perm, perm + nrows, p.begin());
    } else if (HIP_MatLocalityOrder(nrows, rowptr, col, 512, p.data()) != 0) {
        latch(117, "HIP_MatCreateFromCSROrdered: ordering failed");
        return nullptr;
    }
    for (int i = 0; i < nrows; i++) {
        if (p[i] < 0 || p[i] >= nrows || ip[p[i]] != -1) {
            latch(116, "HIP_MatCreateFromCSROrdered: perm is not a permutation (entry %d)", i);
            return nullptr;
        }
        ip[p[i]] = i;
    }
    // row i of P A P^T is the caller's row p[i], its entries in their order,
    // columns renumbered
    std::vector<int> rp((size_t)nrows + 1, 0), cj((size_t)std::max(1, nnz)), nzmap((size_t)std::max(1, nnz));
    std::vector<double> v((size_t)std::max(1, nnz));
    for (int i = 0; i < nrows; i++) {
        const int o = p[i];
        int q = rp[i];
        for (int e = rowptr[o]; e < rowptr[o + 1]; e++, q++) {
            cj[q] = ip[col[e]];
            v[q] = val[e];
            nzmap[e] = q;
        }
        rp[i + 1] = q;
    }
    HIP_MAT *A = HIP_MatCreateFromCSR(nrows, ncols, nnz, rp.data(), cj.data(), v.data());
    if (!A) return nullptr;
    A->perm_h = new int[nrows];
    A->iperm_h = new int[nrows];
    A->nzmap_h = new int[std::max(1, nnz)];
    std::copy(p.begin(), p.end(), A->perm_h);
    std::copy(ip.begin(), ip.end(), A->iperm_h);
    std::copy(nzmap.begin(), nzmap.begin() + std::max(1, nnz), A->nzmap_h);
    return A;
}

extern "C" GCGE_INT HIP_MatGetOrder(const HIP_MAT *mat, GCGE_INT *perm)
{
    if (!mat || mat->magic != kMagicMat || !perm) return -1;
    if (!mat->perm_h) return 0;
    std::copy(mat->perm_h, mat->perm_h + mat->n, perm);
    return 1;
}

extern "C" GCGE_INT HIP_MatGatherKind(const HIP_MAT *mat)
{
    if (!mat || mat->magic != kMagicMat) return -1;
    return mat->gather_rows ? 1 : mat->halo ? 2 : 0;
}

extern "C" void HIP_MatDestroy(HIP_MAT **mat)
{
    if (!mat || !*mat) return;
    HIP_MAT *A = *mat;
    if (A->magic == kMagicMat) {
        sync();
        if (A->rowptr) (void)hipFree(A->rowptr);
        if (A->col) (void)hipFree(A->col);
        if (A->val64) (void)hipFree(A->val64);
        if (A->val32) (void)hipFree(A->val32);
        if (A->rowidx) (void)hipFree(A->rowidx);
        if (A->dinv64) (void)hipFree(A->dinv64);
        if (A->long_rows) (void)hipFree(A->long_rows);
        if (A->dinv32) (void)hipFree(A->dinv32);
        if (A->halo_ptr) (void)hipFree(A->halo_ptr);
        if (A->halo_col) (void)hipFree(A->halo_col);
        if (A->lcol) (void)hipFree(A->lcol);
        delete[] A->perm_h;
        delete[] A->iperm_h;
        delete[] A->nzmap_h;
        g_ctx.dev_bytes -= sizeof(int) * (double)(A->n + 1) + 12.0 * A->nnz + (A->val32 ? 4.0 * A->nnz : 0.0) +
                           (A->rowidx ? 4.0 * A->nnz : 0.0) + A->halo_bytes;
        A->magic = 0;
    }
    delete A;
    *mat = nullptr;
}

extern "C" GCGE_INT HIP_MatNumRows(const HIP_MAT *mat) { return mat ? mat->n : 0; }
extern "C" GCGE_INT HIP_MatNumNonzeros(const HIP_MAT *mat) { return mat ? mat->nnz : 0; }

extern "C" GCGE_INT HIP_MatSetValues(HIP_MAT *A, const GCGE_DOUBLE *val)
{
    if (!A || A->magic != kMagicMat || !val) {
        latch(114, "HIP_MatSetValues: bad arguments");
        return 114;
    }
    std::vector<double> mapped;
    if (A->nzmap_h) {  // the caller's entry order -> the stored order
        mapped.resize((size_t)A->nnz);
        for (int e = 0; e < A->nnz; e++) mapped[A->nzmap_h[e]] = val[e];
        val = mapped.data();
    }
    if (!sync()) return g_err;  // kernels queued on the backend stream may still read the old values
    if (!check(hipMemcpy(A->val64, val, sizeof(double) * A->nnz, hipMemcpyHostToDevice), "H2D val")) return g_err;
    // everything derived from the values follows them: the FP32 copy, the
    // Jacobi diagonal of the preconditioned inner solves, the Gershgorin bound
    if (A->val32) launch_to_f32(A->nnz, A->val64, A->val32);
    if (A->dinv64) launch_diag_inv(A);
    std::vector<int> rp((size_t)A->n + 1);
    if (!check(hipMemcpy(rp.data(), A->rowptr, sizeof(int) * rp.size(), hipMemcpyDeviceToHost), "D2H rowptr"))
        return g_err;
    double rowabs = 0.0;
    for (int i = 0; i < A->n; i++) {
        double sum = 0.0;
        for (int e = rp[i]; e < rp[i + 1]; e++) sum += std::fabs(val[e]);
        rowabs = std::max(rowabs, sum);
    }
    A->row_abs_max = rowabs;
    return g_err;
}

extern "C" GCGE_INT HIP_MatSetPrecision(HIP_MAT *A, GCGE_INT precision)
{
    if (!A || A->magic != kMagicMat || (precision != GCGE_HIP_FP64 && precision != GCGE_HIP_FP32)) {
        latch(115, "HIP_MatSetPrecision: bad arguments");
        return 115;
    }
    A->prec = precision;
    if (precision == GCGE_HIP_FP32) mat_val32(A);
    return g_err;
}

// ====================================================================== multivectors (public)
extern "C" GCGE_INT HIP_MultiVecCreate(void ***mv, GCGE_INT n_vec, HIP_MAT *mat)
{
    if (!mv || !mat || mat->magic != kMagicMat) {
        latch(130, "HIP_MultiVecCreate: bad arguments");
        return 130;
    }
    MV *m = mv_new(mat->n, n_vec, mat->prec, mat);
    *mv = (void **)m;
    return m ? 0 : g_err;
}

extern "C" void HIP_MultiVecDestroy(void ***mv)
{
    if (!mv || !*mv) return;
    mv_free(as_mv(*mv));
    *mv = nullptr;
}

// the caller's row order of a multivector: that of the matrix it was created from
static const int *mv_perm(const MV *m) { return (m->mat && m->mat->perm_h) ? m->mat->perm_h : nullptr; }

template <typename T>
static void host_to_dev_cols(MV *m, int start, int ncols, const double *host, int ldh)
{
    std::vector<T> buf((size_t)m->n);
    const int *perm = mv_perm(m);
    for (int j = 0; j < ncols && !failed(); j++) {
        const double *hc = host + (size_t)j * ldh;
        if (perm)
            for (int64_t i = 0; i < m->n; i++) buf[i] = (T)hc[perm[i]];
        else
            for (int64_t i = 0; i < m->n; i++) buf[i] = (T)hc[i];
        check(hipMemcpy(col_ptr(m, start + j), buf.data(), sizeof(T) * m->n, hipMemcpyHostToDevice), "H2D column");
    }
}

extern "C" GCGE_INT HIP_MultiVecSetFromHost(void **mv, GCGE_INT start, GCGE_INT ncols, const GCGE_DOUBLE *host,
                                            GCGE_INT ldh)
{
    MV *m = as_mv(mv);
    if (!range_ok(m, start, start + ncols, "HIP_MultiVecSetFromHost")) return g_err;
    if (ldh < m->n) {
        latch(131, "HIP_MultiVecSetFromHost: ldh %d < n", ldh);
        return g_err;
    }
    if (!sync()) return g_err;
    if (mv_perm(m)) {
        if (m->prec == GCGE_HIP_FP64) host_to_dev_cols<double>(m, start, ncols, host, ldh);
        else host_to_dev_cols<float>(m, start, ncols, host, ldh);
    } else if (m->prec == GCGE_HIP_FP64) {
        if (ldh == m->ld) {
            check(hipMemcpy(col_ptr(m, start), host, sizeof(double) * (size_t)m->ld * ncols, hipMemcpyHostToDevice),
                  "H2D block");
        } else {
            check(hipMemcpy2D(col_ptr(m, start), sizeof(double) * m->ld, host, sizeof(double) * ldh,
                              sizeof(double) * m->n, ncols, hipMemcpyHostToDevice),
                  "H2D 2D");
        }
    } else {
        host_to_dev_cols<float>(m, start, ncols, host, ldh);
    }
    return g_err;
}

extern "C" GCGE_INT HIP_MultiVecGetToHost(void **mv, GCGE_INT start, GCGE_INT ncols, GCGE_DOUBLE *host,
                                          GCGE_INT ldh)
{
    MV *m = as_mv(mv);
    if (!range_ok(m, start, start + ncols, "HIP_MultiVecGetToHost")) return g_err;
    if (ldh < m->n) {
        latch(132, "HIP_MultiVecGetToHost: ldh %d < n", ldh);
        return g_err;
    }
    if (!sync()) return g_err;
    const int *perm = mv_perm(m);
    if (m->prec == GCGE_HIP_FP64 && !perm) {
        check(hipMemcpy2D(host, sizeof(double) * ldh, col_ptr(m, start), sizeof(double) * m->ld,
                          sizeof(double) * m->n, ncols, hipMemcpyDeviceToHost),
              "D2H 2D");
    } else if (m->prec == GCGE_HIP_FP64) {
        std::vector<double> buf((size_t)m->n);
        for (int j = 0; j < ncols && !failed(); j++) {
            if (!check(hipMemcpy(buf.data(), col_ptr(m, start + j), sizeof(double) * m->n, hipMemcpyDeviceToHost),
                       "D2H column"))
                break;
            double *hc = host + (size_t)j * ldh;
            for (int64_t i = 0; i < m->n; i++) hc[perm[i]] = buf[i];
        }
    } else {
        std::vector<float> buf((size_t)m->n);
        for (int j = 0; j < ncols && !failed(); j++) {
            if (!check(hipMemcpy(buf.data(), col_ptr(m, start + j), sizeof(float) * m->n, hipMemcpyDeviceToHost),
                       "D2H column"))
                break;
            double *hc = host + (size_t)j * ldh;
            if (perm)
                for (int64_t i = 0; i < m->n; i++) hc[perm[i]] = buf[i];
            else
                for (int64_t i = 0; i < m->n; i++) hc[i] = buf[i];
        }
    }
    return g_err;
}

extern "C" GCGE_INT HIP_MultiVecNumCols(void **mv)
{
    MV *m = as_mv(mv);
    return m ? m->ncols : -1;
}

extern "C" GCGE_INT HIP_MultiVecPrecision(void **mv)
{
    MV *m = as_mv(mv);
    return m ? m->prec : -1;
}

// ====================================================================== GCGE_OPS slots
namespace {

MV *vec_mv(void *v, int *col, const char *what)
{
    Col *c = as_col(v);
    if (!c) {
        latch(140, "%s: not a device vector", what);
        return nullptr;
    }
    *col = c->col;
    return c->mv;
}

void s_VecSetRandomValue(void *vec)
{
    int c;
    MV *m = vec_mv(vec, &c, "VecSetRandomValue");
    if (!m) return;
    // rand() / (RAND_MAX + 1) in row order, as CSR_VecSetRandomValue
    std::vector<double> h((size_t)m->n);
    for (int64_t i = 0; i < m->n; i++) h[i] = ((double)rand()) / ((double)RAND_MAX + 1);
    HIP_MultiVecSetFromHost((void **)m, c, 1, h.data(), (int)m->n);
}

void s_MatDotVec(void *Mat, void *x, void *r)
{
    int cx, cr;
    MV *mx = vec_mv(x, &cx, "MatDotVec"), *mr = vec_mv(r, &cr, "MatDotVec");
    HIP_MAT *A = (HIP_MAT *)Mat;
    if (!mx || !mr || !same_rows(mx, mr, "MatDotVec") || !same_prec(mx, mr, "MatDotVec") || !mat_ok(A, mx, "MatDotVec"))
        return;
    launch_spmm(mx->prec, A, col_ptr(mx, cx), mx->ld, col_ptr(mr, cr), mr->ld, 1, nullptr);
}

void s_VecAxpby(GCGE_DOUBLE a, void *x, GCGE_DOUBLE b, void *y)
{
    int cx, cy;
    MV *mx = vec_mv(x, &cx, "VecAxpby"), *my = vec_mv(y, &cy, "VecAxpby");
    if (!mx || !my || !same_rows(mx, my, "VecAxpby") || !same_prec(mx, my, "VecAxpby")) return;
    launch_axpby(mx->prec, mx->n, col_ptr(mx, cx), mx->ld, col_ptr(my, cy), my->ld, 1, &a, &b, nullptr);
}

void s_VecInnerProd(void *x, void *y, GCGE_DOUBLE *v)
{
    int cx, cy;
    MV *mx = vec_mv(x, &cx, "VecInnerProd"), *my = vec_mv(y, &cy, "VecInnerProd");
    if (!mx || !my || !same_rows(mx, my, "VecInnerProd") || !same_prec(mx, my, "VecInnerProd")) return;
    if (!dots(mx->prec, mx->n, col_ptr(mx, cx), mx->ld, col_ptr(my, cy), my->ld, 1, v)) *v = NAN;
}

void s_VecCreateByVec(void **des, void *src)
{
    int c;
    MV *ms = vec_mv(src, &c, "VecCreateByVec");
    *des = nullptr;
    if (!ms) return;
    MV *m = mv_new(ms->n, 1, ms->prec, ms->mat);
    if (!m) return;
    m->cols[0].owner = 1;
    *des = &m->cols[0];
}

void s_VecCreateByMat(void **vec, void *mat)
{
    HIP_MAT *A = (HIP_MAT *)mat;
    *vec = nullptr;
    if (!A || A->magic != kMagicMat) {
        latch(141, "VecCreateByMat: not a HIP_MAT");
        return;
    }
    MV *m = mv_new(A->n, 1, A->prec, A);
    if (!m) return;
    m->cols[0].owner = 1;
    *vec = &m->cols[0];
}

void s_VecDestroy(void **vec)
{
    if (!vec || !*vec) return;
    Col *c = as_col(*vec);
    if (c && c->owner) mv_free(c->mv);
    *vec = nullptr;
}

void s_MultiVecCreateByVec(void ***mv, GCGE_INT n_vec, void *vec, GCGE_OPS *)
{
    int c;
    MV *ms = vec_mv(vec, &c, "MultiVecCreateByVec");
    *mv = nullptr;
    if (!ms) return;
    *mv = (void **)mv_new(ms->n, n_vec, ms->prec, ms->mat);
}

void s_MultiVecCreateByMat(void ***mv, GCGE_INT n_vec, void *mat, GCGE_OPS *)
{
    HIP_MultiVecCreate(mv, n_vec, (HIP_MAT *)mat);
}

void s_MultiVecCreateByMultiVec(void ***mv, GCGE_INT n_vec, void **init, GCGE_OPS *)
{
    MV *ms = as_mv(init);
    *mv = nullptr;
    if (!ms) {
        latch(142, "MultiVecCreateByMultiVec: not a device multivector");
        return;
    }
    *mv = (void **)mv_new(ms->n, n_vec, ms->prec, ms->mat);
}

void s_MultiVecDestroy(void ***mv, GCGE_INT, GCGE_OPS *) { HIP_MultiVecDestroy(mv); }

void s_MultiVecSetRandomValue(void **mv, GCGE_INT start, GCGE_INT n_vec, GCGE_OPS *)
{
    MV *m = as_mv(mv);
    if (!range_ok(m, start, start + n_vec, "MultiVecSetRandomValue")) return;
    for (int i = 0; i < n_vec && !failed(); i++) s_VecSetRandomValue(&m->cols[start + i]);
}

void s_MatDotMultiVec(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    const int k = end[0] - start[0];
    if (!range_ok(mx, start[0], end[0], "MatDotMultiVec") || !range_ok(my, start[1], start[1] + k, "MatDotMultiVec") ||
        !same_rows(mx, my, "MatDotMultiVec") || !same_prec(mx, my, "MatDotMultiVec") ||
        !mat_ok((HIP_MAT *)mat, mx, "MatDotMultiVec"))
        return;
    launch_spmm(mx->prec, (HIP_MAT *)mat, col_ptr(mx, start[0]), mx->ld, col_ptr(my, start[1]), my->ld, k, nullptr);
}

void axpby_cols(double a, MV *mx, int sx, double b, MV *my, int sy, int k)
{
    std::vector<double> av(kMaxCols, a), bv(kMaxCols, b);
    for (int c0 = 0; c0 < k && !failed(); c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        launch_axpby(mx->prec, mx->n, col_ptr(mx, sx + c0), mx->ld, col_ptr(my, sy + c0), my->ld, kc, av.data(),
                     bv.data(), nullptr);
    }
}

void s_MultiVecAxpby(GCGE_DOUBLE a, void **x, GCGE_DOUBLE b, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    const int k = end[0] - start[0];
    if (!range_ok(mx, start[0], end[0], "MultiVecAxpby") || !range_ok(my, start[1], start[1] + k, "MultiVecAxpby") ||
        !same_rows(mx, my, "MultiVecAxpby") || !same_prec(mx, my, "MultiVecAxpby"))
        return;
    axpby_cols(a, mx, start[0], b, my, start[1], k);
}

void s_MultiVecAxpbyColumn(GCGE_DOUBLE a, void **x, GCGE_INT col_x, GCGE_DOUBLE b, void **y, GCGE_INT col_y, GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    if (!range_ok(mx, col_x, col_x + 1, "MultiVecAxpbyColumn") || !range_ok(my, col_y, col_y + 1, "MultiVecAxpbyColumn") ||
        !same_rows(mx, my, "MultiVecAxpbyColumn") || !same_prec(mx, my, "MultiVecAxpbyColumn"))
        return;
    axpby_cols(a, mx, col_x, b, my, col_y, 1);
}

void lincomb_impl(void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_DOUBLE *a, GCGE_INT lda, double alpha,
                  double beta, const char *what)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    const int p = end[0] - start[0], q = end[1] - start[1];
    if (!range_ok(mx, start[0], end[0], what) || !range_ok(my, start[1], end[1], what) || !same_rows(mx, my, what) ||
        !same_prec(mx, my, what))
        return;
    if (p < 1 || lda < p) {
        latch(150, "%s: empty combination or lda %d < %d", what, lda, p);
        return;
    }
    if (mx == my && start[0] < end[1] && start[1] < end[0]) {
        latch(151, "%s: input and output columns overlap", what);
        return;
    }
    lincomb(mx->prec, mx->n, col_ptr(mx, start[0]), mx->ld, p, col_ptr(my, start[1]), my->ld, q, a, lda, alpha, beta);
}

void s_MultiVecLinearComb(void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_DOUBLE *a, GCGE_INT lda, void *,
                          GCGE_INT, GCGE_OPS *)
{
    lincomb_impl(x, y, start, end, a, lda, 1.0, 0.0, "MultiVecLinearComb");
}

void s_MultiVecInnerProd(void **V, void **W, GCGE_DOUBLE *a, char *is_sym, GCGE_INT *start, GCGE_INT *end,
                         GCGE_INT lda, GCGE_OPS *)
{
    MV *mv = as_mv(V), *mw = as_mv(W);
    const bool sym = is_sym && std::strcmp(is_sym, "sym") == 0;
    const int q = end[1] - start[1];
    const int p = sym ? q : end[0] - start[0];
    if (!range_ok(mv, start[0], start[0] + p, "MultiVecInnerProd") ||
        !range_ok(mw, start[1], end[1], "MultiVecInnerProd") || !same_rows(mv, mw, "MultiVecInnerProd") ||
        !same_prec(mv, mw, "MultiVecInnerProd"))
        return;
    if (p == 0 || q == 0) return;
    if (lda < p) {
        latch(152, "MultiVecInnerProd: lda %d < %d", lda, p);
        return;
    }
    if (!sym) {
        gram(mv->prec, mv->n, col_ptr(mv, start[0]), mv->ld, p, col_ptr(mw, start[1]), mw->ld, q, a, lda, false);
        return;
    }
    // "sym": upper triangle computed, then mirrored (gcge_ops.c:206-238)
    std::vector<double> g((size_t)p * q);
    if (!gram(mv->prec, mv->n, col_ptr(mv, start[0]), mv->ld, p, col_ptr(mw, start[1]), mw->ld, q, g.data(), p, true))
        return;
    for (int iw = 0; iw < q; iw++)
        for (int iv = 0; iv <= iw; iv++) a[(size_t)iw * lda + iv] = g[(size_t)iw * p + iv];
    for (int iw = 0; iw < q; iw++)
        for (int iv = iw + 1; iv < q; iv++) a[(size_t)iw * lda + iv] = a[(size_t)iv * lda + iw];
}

void s_MultiVecSwap(void **V1, void **V2, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *)
{
    MV *m1 = as_mv(V1), *m2 = as_mv(V2);
    const int k = end[0] - start[0];
    if (!range_ok(m1, start[0], end[0], "MultiVecSwap") || !range_ok(m2, start[1], start[1] + k, "MultiVecSwap") ||
        !same_rows(m1, m2, "MultiVecSwap") || !same_prec(m1, m2, "MultiVecSwap"))
        return;
    if (m1 == m2 && start[0] == start[1]) return;
    const bool overlap = (m1 == m2) && start[0] < start[1] + k && start[1] < start[0] + k;
    if (overlap) {
        // exchange column by column in order, as the pointer swaps of
        // GCGE_Default_MultiVecSwap (gcge_ops.c:285-300) would
        for (int i = 0; i < k && !failed(); i++)
            launch_swap(m1->prec, col_ptr(m1, start[0] + i), m1->ld, col_ptr(m2, start[1] + i), m2->ld, m1->n, 1);
    } else {
        for (int c0 = 0; c0 < k && !failed(); c0 += 1024)
            launch_swap(m1->prec, col_ptr(m1, start[0] + c0), m1->ld, col_ptr(m2, start[1] + c0), m2->ld, m1->n,
                        std::min(1024, k - c0));
    }
}

void s_MultiVecPrint(void **x, GCGE_INT n)
{
    MV *m = as_mv(x);
    if (!range_ok(m, 0, n, "MultiVecPrint")) return;
    std::vector<double> h((size_t)m->n);
    for (int j = 0; j < n; j++) {
        if (HIP_MultiVecGetToHost(x, j, 1, h.data(), (int)m->n)) return;
        printf("num size = %lld\n", (long long)m->n);
        for (int64_t i = 0; i < m->n; i++) printf("%0.4f\t", h[i]);
        printf("\n");
    }
}

void s_GetVecFromMultiVec(void **V, GCGE_INT j, void **x)
{
    MV *m = as_mv(V);
    if (!range_ok(m, j, j + 1, "GetVecFromMultiVec")) {
        *x = nullptr;
        return;
    }
    *x = &m->cols[j];
}

void s_RestoreVecForMultiVec(void **V, GCGE_INT j, void **x)
{
    MV *m = as_mv(V);
    if (!range_ok(m, j, j + 1, "RestoreVecForMultiVec")) return;
    if (*x != &m->cols[j]) latch(153, "RestoreVecForMultiVec: a foreign vector cannot be placed into a multivector");
}

// CSR_SetDirichletBoundary (app/csr/gcge_app_csr.c:130-183): zero row i of
// the nev vectors when row i of B has no entry above 1e-20 and row i of A
// has exactly one.
void s_SetDirichletBoundary(void **Vecs, GCGE_INT nev, void *Av, void *Bv)
{
    HIP_MAT *A = (HIP_MAT *)Av, *B = (HIP_MAT *)Bv;
    if (!A || !B) return;
    MV *m = as_mv(Vecs);
    if (!range_ok(m, 0, nev, "SetDirichletBoundary") || !mat_ok(A, m, "SetDirichletBoundary") ||
        !mat_ok(B, m, "SetDirichletBoundary"))
        return;
    const int n = A->n;
    std::vector<int> arp(n + 1), brp(n + 1);
    std::vector<double> av(A->nnz), bv(B->nnz);
    if (!sync()) return;
    if (!check(hipMemcpy(arp.data(), A->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(brp.data(), B->rowptr, sizeof(int) * (n + 1), hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(av.data(), A->val64, sizeof(double) * A->nnz, hipMemcpyDeviceToHost), "D2H") ||
        !check(hipMemcpy(bv.data(), B->val64, sizeof(double) * B->nnz, hipMemcpyDeviceToHost), "D2H"))
        return;
    std::vector<unsigned char> mask(n, 0);
    int count = 0;
    for (int i = 0; i < n; i++) {
        bool bzero = true;
        for (int e = brp[i]; e < brp[i + 1]; e++)
            if (std::fabs(bv[e]) > 1e-20) {
                bzero = false;
                break;
            }
        if (!bzero) continue;
        int na = 0;
        for (int e = arp[i]; e < arp[i + 1]; e++)
            if (std::fabs(av[e]) > 1e-20) na++;
        if (na == 1) {
            mask[i] = 1;
            count++;
        }
    }
    if (!count) return;
    unsigned char *dmask = (unsigned char *)scratch(n);
    if (!dmask) return;
    if (!check(hipMemcpy(dmask, mask.data(), n, hipMemcpyHostToDevice), "H2D mask")) return;
    launch_zero_rows(m->prec, m->data, m->ld, m->n, nev, dmask);
    sync();
}

// ---- block-kernel slots ----
void s_MultiVecAxpbyDiag(const GCGE_DOUBLE *a, void **x, const GCGE_DOUBLE *b, void **y, GCGE_INT *start,
                         GCGE_INT *end, GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    const int k = end[0] - start[0];
    if (!range_ok(mx, start[0], end[0], "MultiVecAxpbyDiag") || !range_ok(my, start[1], start[1] + k, "MultiVecAxpbyDiag") ||
        !same_rows(mx, my, "MultiVecAxpbyDiag") || !same_prec(mx, my, "MultiVecAxpbyDiag"))
        return;
    for (int c0 = 0; c0 < k && !failed(); c0 += kMaxCols)
        launch_axpby(mx->prec, mx->n, col_ptr(mx, start[0] + c0), mx->ld, col_ptr(my, start[1] + c0), my->ld,
                     std::min(kMaxCols, k - c0), a + c0, b + c0, nullptr);
}

void s_MultiVecInnerProdDiag(void **x, void **y, GCGE_DOUBLE *out, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    const int k = end[0] - start[0];
    if (!range_ok(mx, start[0], end[0], "MultiVecInnerProdDiag") ||
        !range_ok(my, start[1], start[1] + k, "MultiVecInnerProdDiag") || !same_rows(mx, my, "MultiVecInnerProdDiag") ||
        !same_prec(mx, my, "MultiVecInnerProdDiag"))
        return;
    if (!dots(mx->prec, mx->n, col_ptr(mx, start[0]), mx->ld, col_ptr(my, start[1]), my->ld, k, out))
        for (int j = 0; j < k; j++) out[j] = NAN;
}

void s_MultiVecLinearCombAxpby(void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_DOUBLE *a, GCGE_INT lda,
                               GCGE_DOUBLE alpha, GCGE_DOUBLE beta, GCGE_OPS *)
{
    lincomb_impl(x, y, start, end, a, lda, alpha, beta, "MultiVecLinearCombAxpby");
}

void s_MatDotMultiVecInnerProdDiag(void *mat, void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_DOUBLE *out,
                                   GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    const int k = end[0] - start[0];
    if (!range_ok(mx, start[0], end[0], "MatDotMultiVecInnerProdDiag") ||
        !range_ok(my, start[1], start[1] + k, "MatDotMultiVecInnerProdDiag") ||
        !same_rows(mx, my, "MatDotMultiVecInnerProdDiag") || !same_prec(mx, my, "MatDotMultiVecInnerProdDiag") ||
        !mat_ok((HIP_MAT *)mat, mx, "MatDotMultiVecInnerProdDiag"))
        return;
    if (!spmm_dot(mx->prec, (HIP_MAT *)mat, col_ptr(mx, start[0]), mx->ld, col_ptr(my, start[1]), my->ld, k, out))
        for (int j = 0; j < k; j++) out[j] = NAN;
}

void s_BlockCGUpdate(void **x, GCGE_INT sx, void **p, void **w, void **r, GCGE_INT k, const GCGE_DOUBLE *alpha,
                     GCGE_DOUBLE *rr, GCGE_OPS *)
{
    MV *mx = as_mv(x), *mp = as_mv(p), *mw = as_mv(w), *mr = as_mv(r);
    if (!range_ok(mx, sx, sx + k, "BlockCGUpdate") || !range_ok(mp, 0, k, "BlockCGUpdate") ||
        !range_ok(mw, 0, k, "BlockCGUpdate") || !range_ok(mr, 0, k, "BlockCGUpdate") ||
        !same_rows(mx, mp, "BlockCGUpdate") || !same_rows(mx, mw, "BlockCGUpdate") ||
        !same_rows(mx, mr, "BlockCGUpdate") || !same_prec(mx, mp, "BlockCGUpdate") ||
        !same_prec(mx, mw, "BlockCGUpdate") || !same_prec(mx, mr, "BlockCGUpdate"))
        return;
    if (!cg_update(mx->prec, mx->n, col_ptr(mx, sx), mx->ld, mp->data, mp->ld, mw->data, mw->ld, mr->data, mr->ld, k,
                   alpha, rr))
        for (int j = 0; j < k; j++) rr[j] = NAN;
}

void s_BlockCGResidualUpdate(void **w, void **r, GCGE_INT k, const GCGE_DOUBLE *alpha, GCGE_DOUBLE *rr, GCGE_OPS *)
{
    MV *mw = as_mv(w), *mr = as_mv(r);
    if (!range_ok(mw, 0, k, "BlockCGResidualUpdate") || !range_ok(mr, 0, k, "BlockCGResidualUpdate") ||
        !same_rows(mw, mr, "BlockCGResidualUpdate") || !same_prec(mw, mr, "BlockCGResidualUpdate"))
        return;
    if (!cg_resid(mr->prec, mr->n, mw->data, mw->ld, mr->data, mr->ld, k, alpha, rr))
        for (int j = 0; j < k; j++) rr[j] = NAN;
}

void s_BlockPCGResidualUpdate(void *Mat, void **w, void **r, GCGE_INT k, const GCGE_DOUBLE *alpha, GCGE_DOUBLE *rz,
                              GCGE_DOUBLE *rr, GCGE_OPS *)
{
    HIP_MAT *A = (HIP_MAT *)Mat;
    MV *mw = as_mv(w), *mr = as_mv(r);
    const char *who = "BlockPCGResidualUpdate";
    if (!range_ok(mw, 0, k, who) || !range_ok(mr, 0, k, who) || !same_rows(mw, mr, who) || !same_prec(mw, mr, who) ||
        !mat_ok(A, mr, who))
        return;
    const void *dinv = mat_dinv(A, mr->prec);
    if (!dinv || !pcg_resid(mr->prec, mr->n, mw->data, mw->ld, mr->data, mr->ld, dinv, k, alpha, rz, rr))
        for (int j = 0; j < k; j++) rz[j] = rr[j] = NAN;
}

void s_BlockPCGSearchUpdate(void *Mat, void **x, GCGE_INT sx, void **p, void **r, GCGE_INT kx, GCGE_INT kp,
                            const GCGE_DOUBLE *alpha, const GCGE_DOUBLE *beta, GCGE_OPS *)
{
    HIP_MAT *A = (HIP_MAT *)Mat;
    MV *mx = as_mv(x), *mp = as_mv(p), *mr = as_mv(r);
    const char *who = "BlockPCGSearchUpdate";
    if (kp < 0 || kx < 0) {
        latch(172, "%s: negative column count (kx %d, kp %d)", who, kx, kp);
        return;
    }
    const int k = std::max(kx, kp);
    if (!range_ok(mx, sx, sx + kx, who) || !range_ok(mp, 0, k, who) || !range_ok(mr, 0, kp, who) ||
        !same_rows(mx, mp, who) || !same_rows(mx, mr, who) || !same_prec(mx, mp, who) || !same_prec(mx, mr, who) ||
        !mat_ok(A, mr, who))
        return;
    const void *dinv = mat_dinv(A, mr->prec);
    if (!dinv) return;
    launch_cg_search(mx->prec, mx->n, col_ptr(mx, sx), mx->ld, mp->data, mp->ld, mr->data, mr->ld, kx, kp, alpha,
                     beta, dinv);
}

void s_BlockCGSearchUpdate(void **x, GCGE_INT sx, void **p, void **r, GCGE_INT kx, GCGE_INT kp,
                           const GCGE_DOUBLE *alpha, const GCGE_DOUBLE *beta, GCGE_OPS *)
{
    MV *mx = as_mv(x), *mp = as_mv(p), *mr = as_mv(r);
    if (kp < 0 || kx < 0) {
        latch(172, "BlockCGSearchUpdate: negative column count (kx %d, kp %d)", kx, kp);
        return;
    }
    const int k = std::max(kx, kp);
    if (!range_ok(mx, sx, sx + kx, "BlockCGSearchUpdate") || !range_ok(mp, 0, k, "BlockCGSearchUpdate") ||
        !range_ok(mr, 0, kp, "BlockCGSearchUpdate") || !same_rows(mx, mp, "BlockCGSearchUpdate") ||
        !same_rows(mx, mr, "BlockCGSearchUpdate") || !same_prec(mx, mp, "BlockCGSearchUpdate") ||
        !same_prec(mx, mr, "BlockCGSearchUpdate"))
        return;
    launch_cg_search(mx->prec, mx->n, col_ptr(mx, sx), mx->ld, mp->data, mp->ld, mr->data, mr->ld, kx, kp, alpha,
                     beta);
}

void s_MultiVecResidualNorms(void *A, void *B, void **X, const GCGE_INT *idx, GCGE_INT k, const GCGE_DOUBLE *lambda,
                             GCGE_DOUBLE *rnorm, GCGE_DOUBLE *xnorm, GCGE_DOUBLE *bxnorm, GCGE_OPS *)
{
    MV *mx = as_mv(X);
    if (!mx) {
        latch(160, "MultiVecResidualNorms: not a device multivector");
        return;
    }
    for (int j = 0; j < k; j++)
        if (idx[j] < 0 || idx[j] >= mx->ncols) {
            latch(161, "MultiVecResidualNorms: column %d outside [0,%d)", idx[j], mx->ncols);
            return;
        }
    if (!mat_ok((HIP_MAT *)A, mx, "MultiVecResidualNorms") || (B && !mat_ok((HIP_MAT *)B, mx, "MultiVecResidualNorms")))
        return;
    if (!residual_norms(mx->prec, (HIP_MAT *)A, (HIP_MAT *)B, mx->data, mx->ld, idx, k, lambda, rnorm, xnorm, bxnorm))
        for (int j = 0; j < k; j++) rnorm[j] = NAN;
}

void s_MultiVecGather(void **x, const GCGE_INT *idx, GCGE_INT k, void **y, GCGE_INT ystart, GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    if (!range_ok(my, ystart, ystart + k, "MultiVecGather") || !mx || !same_rows(mx, my, "MultiVecGather") ||
        !same_prec(mx, my, "MultiVecGather"))
        return;
    for (int j = 0; j < k; j++) {
        if (idx[j] < 0 || idx[j] >= mx->ncols) {
            latch(162, "MultiVecGather: column %d outside [0,%d)", idx[j], mx->ncols);
            return;
        }
        if (mx == my && idx[j] >= ystart && idx[j] < ystart + k) {
            latch(163, "MultiVecGather: source column %d inside the destination range", idx[j]);
            return;
        }
    }
    std::vector<double> one(kMaxCols, 1.0), zero(kMaxCols, 0.0);
    for (int c0 = 0; c0 < k && !failed(); c0 += kMaxCols)
        launch_axpby(mx->prec, mx->n, mx->data, mx->ld, col_ptr(my, ystart + c0), my->ld, std::min(kMaxCols, k - c0),
                     one.data(), zero.data(), idx + c0);
}

void s_MultiVecCopy(void **x, void **y, GCGE_INT *start, GCGE_INT *end, GCGE_OPS *)
{
    MV *mx = as_mv(x), *my = as_mv(y);
    const int k = end[0] - start[0];
    if (!range_ok(mx, start[0], end[0], "MultiVecCopy") || !range_ok(my, start[1], start[1] + k, "MultiVecCopy") ||
        !same_rows(mx, my, "MultiVecCopy"))
        return;
    if (mx == my && start[0] < start[1] + k && start[1] < start[0] + k) {
        latch(164, "MultiVecCopy: overlapping ranges");
        return;
    }
    for (int c0 = 0; c0 < k && !failed(); c0 += 1024)
        launch_copy(mx->prec, col_ptr(mx, start[0] + c0), mx->ld, my->prec, col_ptr(my, start[1] + c0), my->ld, mx->n,
                    std::min(1024, k - c0));
}

void s_MultiVecCreateByMatPrecision(void ***mv, GCGE_INT n_vec, void *mat, GCGE_INT precision, GCGE_OPS *)
{
    HIP_MAT *A = (HIP_MAT *)mat;
    *mv = nullptr;
    if (!A || A->magic != kMagicMat || (precision != GCGE_HIP_FP64 && precision != GCGE_HIP_FP32)) {
        latch(143, "MultiVecCreateByMatPrecision: bad arguments");
        return;
    }
    if (precision == GCGE_HIP_FP32 && !mat_val32(A)) return;
    *mv = (void **)mv_new(A->n, n_vec, precision, A);
}

GCGE_INT s_MultiVecNumRows(void **x)
{
    MV *m = as_mv(x);
    return m ? (GCGE_INT)m->n : -1;
}

}  // namespace

// dense slots live in the host core (gcge_dense.cpp)
extern "C" void GCGE_SetDenseOps(GCGE_OPS *ops);

extern "C" void GCGE_HIP_SetOps(GCGE_OPS *ops)
{
    ensure_ctx();
    ops->VecSetRandomValue = s_VecSetRandomValue;
    ops->MatDotVec = s_MatDotVec;
    ops->VecAxpby = s_VecAxpby;
    ops->VecInnerProd = s_VecInnerProd;
    ops->VecLocalInnerProd = s_VecInnerProd;
    ops->VecCreateByVec = s_VecCreateByVec;
    ops->VecCreateByMat = s_VecCreateByMat;
    ops->VecDestroy = s_VecDestroy;
    ops->MultiVecCreateByVec = s_MultiVecCreateByVec;
    ops->MultiVecCreateByMat = s_MultiVecCreateByMat;
    ops->MultiVecCreateByMultiVec = s_MultiVecCreateByMultiVec;
    ops->MultiVecDestroy = s_MultiVecDestroy;
    ops->MultiVecSetRandomValue = s_MultiVecSetRandomValue;
    ops->MatDotMultiVec = s_MatDotMultiVec;
    ops->MultiVecAxpby = s_MultiVecAxpby;
    ops->MultiVecAxpbyColumn = s_MultiVecAxpbyColumn;
    ops->MultiVecLinearComb = s_MultiVecLinearComb;
    ops->MultiVecInnerProd = s_MultiVecInnerProd;
    ops->MultiVecSwap = s_MultiVecSwap;
    ops->MultiVecPrint = s_MultiVecPrint;
    ops->GetVecFromMultiVec = s_GetVecFromMultiVec;
    ops->RestoreVecForMultiVec = s_RestoreVecForMultiVec;
    ops->SetDirichletBoundary = s_SetDirichletBoundary;
    ops->MultiVecAxpbyDiag = s_MultiVecAxpbyDiag;
    ops->MultiVecInnerProdDiag = s_MultiVecInnerProdDiag;
    ops->MultiVecLinearCombAxpby = s_MultiVecLinearCombAxpby;
    ops->MatDotMultiVecInnerProdDiag = s_MatDotMultiVecInnerProdDiag;
    ops->BlockCGUpdate = s_BlockCGUpdate;
    ops->BlockCGResidualUpdate = s_BlockCGResidualUpdate;
    ops->BlockCGSearchUpdate = s_BlockCGSearchUpdate;
    ops->BlockPCGResidualUpdate = s_BlockPCGResidualUpdate;
    ops->BlockPCGSearchUpdate = s_BlockPCGSearchUpdate;
    ops->MultiVecResidualNorms = s_MultiVecResidualNorms;
    ops->MultiVecGather = s_MultiVecGather;
    ops->MultiVecCopy = s_MultiVecCopy;
    ops->MultiVecNumRows = s_MultiVecNumRows;
    ops->MultiVecCreateByMatPrecision = s_MultiVecCreateByMatPrecision;
    GCGE_SetDenseOps(ops);
}

extern "C" void GCGE_SOLVER_SetHIPOps(GCGE_SOLVER *solver) { GCGE_HIP_SetOps(solver->ops); }

// GCGE_CSR_Solver_Init counterpart (app/csr/gcge_app_csr.c:92-124)
extern "C" GCGE_SOLVER *GCGE_HIP_Solver_Init(HIP_MAT *A, HIP_MAT *B, GCGE_INT num_eigenvalues, GCGE_INT argc,
                                             char *argv[])
{
    if (!A || A->magic != kMagicMat || (B && (B->magic != kMagicMat || B->n != A->n))) {
        latch(170, "GCGE_HIP_Solver_Init: bad matrices");
        return nullptr;
    }
    if (B && (!A->perm_h != !B->perm_h || (A->perm_h && !std::equal(A->perm_h, A->perm_h + A->n, B->perm_h)))) {
        latch(171, "GCGE_HIP_Solver_Init: A and B were created with different row orders");
        return nullptr;
    }
    GCGE_SOLVER *solver;
    GCGE_SOLVER_Create(&solver);
    if (num_eigenvalues != -1) solver->para->nev = num_eigenvalues;
    GCGE_PARA_SetFromCommandLine(solver->para, argc, argv);
    const int nev = solver->para->nev;
    solver->eval = (double *)calloc(nev, sizeof(double));
    void **evec = nullptr;
    HIP_MultiVecCreate(&evec, nev, A);
    GCGE_SOLVER_SetMatA(solver, A);
    if (B) GCGE_SOLVER_SetMatB(solver, B);
    GCGE_SOLVER_SetHIPOps(solver);
    GCGE_SOLVER_SetEigenvalues(solver, solver->eval);
    GCGE_SOLVER_SetEigenvectors(solver, evec);
    GCGE_SOLVER_Setup(solver);
    return solver;
}

extern "C" GCGE_INT GCGE_HIP_GetLastError(char *msg, GCGE_INT len)
{
    if (msg && len > 0) {
        std::strncpy(msg, g_msg, (size_t)len - 1);
        msg[len - 1] = 0;
    }
    return g_err;
}

extern "C" void GCGE_HIP_ClearError(void)
{
    g_err = 0;
    g_msg[0] = 0;
}

extern "C" GCGE_INT GCGE_HIP_Synchronize(void)
{
    sync();
    return g_err;
}

extern "C" double GCGE_HIP_DeviceBytes(void) { return g_ctx.dev_bytes; }

extern "C" void GCGE_HIP_SetProfiling(GCGE_INT on)
{
    sync();
    g_prof = on != 0;
}

extern "C" void GCGE_HIP_ResetKernelStats(void)
{
    sync();
    for (auto &st : g_stats) st = ProfStat();
}

extern "C" GCGE_INT GCGE_HIP_GetKernelStats(GCGE_INT kind, char *name, GCGE_INT len, double *count, double *ms,
                                            double *bytes, double *flops)
{
    if (kind < 0 || kind >= K_NUM) return -1;
    sync();
    if (name && len > 0) {
        std::strncpy(name, kclass_name(kind), (size_t)len - 1);
        name[len - 1] = 0;
    }
    if (count) *count = (double)g_stats[kind].count;
    if (ms) *ms = g_stats[kind].ms;
    if (bytes) *bytes = g_stats[kind].bytes;
    if (flops) *flops = g_stats[kind].flops;
    return K_NUM;
}

extern "C" GCGE_INT GCGE_HIP_AbiLayout(GCGE_INT *out, GCGE_INT n)
{
    const GCGE_INT v[] = {(GCGE_INT)sizeof(GCGE_PARA),
                          (GCGE_INT)offsetof(GCGE_PARA, ev_tol),
                          (GCGE_INT)offsetof(GCGE_PARA, orth_para),
                          (GCGE_INT)offsetof(GCGE_PARA, cg_rate),
                          (GCGE_INT)offsetof(GCGE_PARA, res),
                          (GCGE_INT)offsetof(GCGE_PARA, stat_para),
                          (GCGE_INT)offsetof(GCGE_PARA, opt_allgatherv),
                          (GCGE_INT)offsetof(GCGE_PARA, stagnation_ratio),
                          (GCGE_INT)sizeof(GCGE_ORTH_PARA),
                          (GCGE_INT)sizeof(GCGE_WORKSPACE),
                          (GCGE_INT)offsetof(GCGE_WORKSPACE, unlock),
                          (GCGE_INT)offsetof(GCGE_WORKSPACE, eval),
                          (GCGE_INT)sizeof(GCGE_SOLVER),
                          (GCGE_INT)sizeof(GCGE_OPS),
                          (GCGE_INT)offsetof(GCGE_OPS, ArrayScale),
                          (GCGE_INT)offsetof(GCGE_OPS, MultiVecAxpbyDiag),
                          (GCGE_INT)offsetof(GCGE_PARA, lock_converged),
                          (GCGE_INT)sizeof(GCGE_HIP_MIXED_OPTS),
                          (GCGE_INT)offsetof(GCGE_HIP_MIXED_OPTS, stagnation_start),
                          (GCGE_INT)sizeof(GCGE_HIP_MIXED_INFO)};
    const GCGE_INT m = (GCGE_INT)(sizeof(v) / sizeof(v[0]));
    for (GCGE_INT i = 0; i < m && i < n; i++) out[i] = v[i];
    return m < n ? m : n;
}

// ---- eigenvalue gradient (adjoint) ----
namespace {

// A Jacobian of a matrix's stored values: nparam rows, jrowptr ascending,
// every jnz a stored-entry index of the matrix.
bool jacobian_ok(const HIP_MAT *M, GCGE_INT nparam, const GCGE_INT *jrowptr, const GCGE_INT *jnz,
                 const GCGE_DOUBLE *jval, const char *which)
{
    if (!jrowptr || (jrowptr[nparam] > 0 && (!jnz || !jval)) || jrowptr[0] != 0) {
        latch(180, "GCGE_HIP_EigGradient: bad %s Jacobian arrays", which);
        return false;
    }
    for (int p = 0; p < nparam; p++)
        if (jrowptr[p + 1] < jrowptr[p]) {
            latch(181, "GCGE_HIP_EigGradient: %s jrowptr decreases at %d", which, p);
            return false;
        }
    for (int t = 0; t < jrowptr[nparam]; t++)
        if (jnz[t] < 0 || jnz[t] >= M->nnz) {
            latch(182, "GCGE_HIP_EigGradient: %s jnz[%d]=%d outside [0,%d)", which, t, jnz[t], M->nnz);
            return false;
        }
    return true;
}

// grad (device, nparam) = scale * J^T g_M + (accumulate ? grad : 0), where g_M[e] =
// sum_i wgt_i x_i[row_e] x_i[col_e] over the stored entries of M
bool gradient_term(HIP_MAT *M, MV *mx, GCGE_INT k, const double *wgt, GCGE_INT nparam, const GCGE_INT *jrowptr,
                   const GCGE_INT *jnz, const GCGE_DOUBLE *jval, double scale, bool accumulate, double *dgrad)
{
    const int nj = jrowptr[nparam];
    std::vector<int> mapped;
    if (M->nzmap_h && nj > 0) {  // the caller's entry indices -> the stored order
        mapped.resize((size_t)nj);
        for (int t = 0; t < nj; t++) mapped[t] = M->nzmap_h[jnz[t]];
        jnz = mapped.data();
    }
    double *g = nullptr, *dval = nullptr;
    int *drp = nullptr, *dnz = nullptr;
    bool ok = check(hipMalloc((void **)&g, sizeof(double) * std::max(1, M->nnz)), "hipMalloc g") &&
              check(hipMalloc((void **)&drp, sizeof(int) * (nparam + 1)), "hipMalloc jrowptr") &&
              check(hipMalloc((void **)&dnz, sizeof(int) * std::max(1, nj)), "hipMalloc jnz") &&
              check(hipMalloc((void **)&dval, sizeof(double) * std::max(1, nj)), "hipMalloc jval");
    if (ok)
        ok = check(hipMemcpy(drp, jrowptr, sizeof(int) * (nparam + 1), hipMemcpyHostToDevice), "H2D") &&
             (nj == 0 || check(hipMemcpy(dnz, jnz, sizeof(int) * nj, hipMemcpyHostToDevice), "H2D")) &&
             (nj == 0 || check(hipMemcpy(dval, jval, sizeof(double) * nj, hipMemcpyHostToDevice), "H2D"));
    if (ok) ok = sampled_outer(mx->prec, M, mx->data, mx->ld, k, wgt, g);
    if (ok) {
        launch_csr_gather_sum(nparam, drp, dnz, dval, g, dgrad, scale, accumulate);
        ok = sync();
    }
    sync();
    if (g) (void)hipFree(g);
    if (drp) (void)hipFree(drp);
    if (dnz) (void)hipFree(dnz);
    if (dval) (void)hipFree(dval);
    return ok;
}

}  // namespace

extern "C" GCGE_INT GCGE_HIP_EigGradientGeneralized(HIP_MAT *A, HIP_MAT *B, void **X, GCGE_INT k,
                                                    const GCGE_DOUBLE *lambda, const GCGE_DOUBLE *w, GCGE_INT nparam,
                                                    const GCGE_INT *ja_rowptr, const GCGE_INT *ja_nz,
                                                    const GCGE_DOUBLE *ja_val, const GCGE_INT *jb_rowptr,
                                                    const GCGE_INT *jb_nz, const GCGE_DOUBLE *jb_val, GCGE_DOUBLE *grad)
{
    MV *mx = as_mv(X);
    if (!A || A->magic != kMagicMat || !range_ok(mx, 0, k, "GCGE_HIP_EigGradient") ||
        !mat_ok(A, mx, "GCGE_HIP_EigGradient") || nparam < 0 || !grad || !w ||
        (B && (B->magic != kMagicMat || B->n != A->n)) || (B && jb_rowptr && !lambda)) {
        latch(180, "GCGE_HIP_EigGradient: bad arguments");
        return g_err;
    }
    if (!jacobian_ok(A, nparam, ja_rowptr, ja_nz, ja_val, "A")) return g_err;
    const bool bterm = B && jb_rowptr;
    if (bterm && !jacobian_ok(B, nparam, jb_rowptr, jb_nz, jb_val, "B")) return g_err;
    double *dgrad = nullptr;
    if (!check(hipMalloc((void **)&dgrad, sizeof(double) * std::max(1, nparam)), "hipMalloc grad")) return g_err;
    // x_i^T (dA/dtheta) x_i weighted by w_i, then - lambda_i x_i^T (dB/dtheta) x_i
    bool ok = gradient_term(A, mx, k, w, nparam, ja_rowptr, ja_nz, ja_val, 1.0, false, dgrad);
    if (ok && bterm) {
        std::vector<double> wl(k);
        for (int i = 0; i < k; i++) wl[i] = w[i] * lambda[i];
        ok = gradient_term(B, mx, k, wl.data(), nparam, jb_rowptr, jb_nz, jb_val, -1.0, true, dgrad);
    }
    if (ok) check(hipMemcpy(grad, dgrad, sizeof(double) * nparam, hipMemcpyDeviceToHost), "D2H grad");
    (void)hipFree(dgrad);
    return g_err;
}

extern "C" GCGE_INT GCGE_HIP_EigGradient(HIP_MAT *A, void **X, GCGE_INT k, const GCGE_DOUBLE *w, GCGE_INT nparam,
                                         const GCGE_INT *jrowptr, const GCGE_INT *jnz, const GCGE_DOUBLE *jval,
                                         GCGE_DOUBLE *grad)
{
    return GCGE_HIP_EigGradientGeneralized(A, nullptr, X, k, nullptr, w, nparam, jrowptr, jnz, jval, nullptr, nullptr,
                                           nullptr, grad);
}
