// hip_kernels.hip -- device kernels of the gfx950 GCGE backend.
//
// Compiled with -ffp-contract=off: element-wise kernels (SpMM, axpby, CG
// update, residual) round every product and sum separately, in the order of
// the reference's CPU loops (CSR_MatDotVec and CSR_VecAxpby,
// app/csr/gcge_app_csr.c:297-365), so a row of A x or an axpby is
// bit-identical to the CPU result.  Reductions (dots, Gram) use explicit
// fma and a fixed two-stage order: deterministic run to run, but not the
// CPU's sequential summation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <climits>
#include <vector>

#include "hip_common.h"

namespace gh {

namespace {

constexpr int kBlock = 256;
constexpr int kSpmmRows = 256;   // rows per SpMM workgroup (one row per thread)
constexpr int kSpmmCap = 2048;   // staged nonzeros per SpMM workgroup

struct ColScalars {
    double a[kMaxCols];
    double b[kMaxCols];
    int idx[kMaxCols];
    int use_idx;
};

// y = a x + b y with the branch rules of CSR_VecAxpby (gcge_app_csr.c:330-365)
template <typename T>
__device__ __forceinline__ T axpby_el(T a, T x, T b, T y)
{
    if (a == T(0)) return b * y;
    if (b == T(0)) return (a == T(1)) ? x : a * x;
    return a * x + b * y;
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// sum over a 256-thread block; result valid in thread 0
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

int grid_1d(int64_t n, int cap)
{
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b > cap) b = cap;
    if (b < 1) b = 1;
    return (int)b;
}

// ---------------------------------------------------------------- axpby
template <typename T>
__global__ __launch_bounds__(kBlock) void k_axpby(int64_t n, const T *__restrict__ x, int64_t ldx,
                                                  T *__restrict__ y, int64_t ldy, ColScalars s)
{
    const int j = blockIdx.y;
    const int64_t xc = s.use_idx ? (int64_t)s.idx[j] : (int64_t)j;
    const T *xj = x + xc * ldx;
    T *yj = y + (int64_t)j * ldy;
    const T a = (T)s.a[j], b = (T)s.b[j];
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    if (a == T(0)) {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) yj[i] = b * yj[i];
    } else if (b == T(0)) {
        if (a == T(1)) {
            for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) yj[i] = xj[i];
        } else {
            for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) yj[i] = a * xj[i];
        }
    } else {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
            yj[i] = a * xj[i] + b * yj[i];
    }
}

// ---------------------------------------------------------------- dots
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dots(int64_t n, const T *__restrict__ x, int64_t ldx,
                                                 const T *__restrict__ y, int64_t ldy,
                                                 double *__restrict__ partial)
{
    __shared__ double sh[4];
    const int j = blockIdx.y;
    const T *xj = x + (int64_t)j * ldx;
    const T *yj = y + (int64_t)j * ldy;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        acc = fma((double)xj[i], (double)yj[i], acc);
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = s;
}

// out[j] = sum_b partial[j * nb + b], one block per j, fixed order
__global__ __launch_bounds__(kBlock) void k_reduce_rows(const double *__restrict__ partial, int nb,
                                                        double *__restrict__ out)
{
    __shared__ double sh[4];
    const int j = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) acc += partial[(int64_t)j * nb + b];
    const double s = block_sum(acc, sh);
    if (threadIdx.x == 0) out[j] = s;
}

// ---------------------------------------------------------------- copy / swap / rows
template <typename TX, typename TY>
__global__ __launch_bounds__(kBlock) void k_copy(int64_t n, const TX *__restrict__ x, int64_t ldx,
                                                 TY *__restrict__ y, int64_t ldy)
{
    const int j = blockIdx.y;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        y[(int64_t)j * ldy + i] = (TY)x[(int64_t)j * ldx + i];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_swap(int64_t n, T *__restrict__ x, int64_t ldx,
                                                 T *__restrict__ y, int64_t ldy)
{
    const int j = blockIdx.y;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        T t = x[(int64_t)j * ldx + i];
        x[(int64_t)j * ldx + i] = y[(int64_t)j * ldy + i];
        y[(int64_t)j * ldy + i] = t;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero_rows(int64_t n, T *__restrict__ x, int64_t ldx,
                                                      const unsigned char *__restrict__ mask)
{
    const int j = blockIdx.y;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        if (mask[i]) x[(int64_t)j * ldx + i] = T(0);
}

// ---------------------------------------------------------------- SpMM
// Column sums of d[0..KC) over a wavefront by recursive halving: at level j
// (partner lane ^ (32 >> j)) each lane keeps half of its columns (the upper
// half if its bit is set) and adds the partner's copy of them; after log2 KC
// levels it holds one column, lane >> (6 - log2 KC), which the remaining
// levels sum over the lanes that share it.  7 exchanges for 4 columns, 10 for
// 8, instead of 6 per column; the order is fixed, so the sums are
// reproducible.
template <int KC>
__device__ __forceinline__ double wave_colsum(double (&d)[KC], int lane)
{
    static_assert(KC >= 1 && KC <= 64 && (KC & (KC - 1)) == 0, "KC is a power of two");
    int m = KC;
#pragma unroll
    for (int mask = 32; mask >= 1; mask >>= 1) {
        if (m > 1) {
            const bool hi = (lane & mask) != 0;
            const int h = m / 2;
#pragma unroll
            for (int i = 0; i < KC / 2; i++) {
                if (i < h) {
                    const double send = hi ? d[i] : d[i + h];
                    const double keep = hi ? d[i + h] : d[i];
                    d[i] = keep + __shfl_xor(send, mask, 64);
                }
            }
            m = h;
        } else {
            d[0] = d[0] + __shfl_xor(d[0], mask, 64);
        }
    }
    return d[0];
}

// One thread per row; the workgroup's nonzeros are staged in LDS when they
// fit; columns are processed KC at a time.  Each row sum starts at 0 and
// adds val*x in CSR order, as CSR_MatDotVec does; the nonzeros go in batches
// of kSpmmBatch whose column indices, values and X gathers are all issued
// before the first product is added, so a row costs one memory round trip
// per batch rather than one per nonzero.  With DOT, the workgroup also
// writes partial[j * nblk + blk] = sum_r X(r,j) Y(r,j) (wavefront sums by
// wave_colsum, combined in LDS in a fixed order) at partial[j * pstride +
// blk].  Rows with more than long_thr entries are left to k_spmm_long.
// STREAM (used for FP32): row blocks are renumbered so that each XCD walks a
// contiguous range of rows (the plane neighbours of a stencil stay in its
// L2), and the matrix and Y, each touched once, bypass the caches
// (non-temporal) to leave L2 to X.
constexpr int kSpmmBatch = 4;

template <typename T, bool DOT, int KC, bool STREAM>
__global__ __launch_bounds__(kSpmmRows) void k_spmm(int n, const int *__restrict__ rowptr,
                                                    const int *__restrict__ col,
                                                    const T *__restrict__ val,
                                                    const T *__restrict__ X, int64_t ldx,
                                                    T *__restrict__ Y, int64_t ldy, int k,
                                                    ColScalars g, double *__restrict__ partial, int nblk,
                                                    int long_thr, int pstride)
{
    __shared__ int s_col[kSpmmCap];
    __shared__ T s_val[kSpmmCap];
    __shared__ double sdot[DOT ? kMaxCols : 1][4];
    int blk = blockIdx.x;
    if (STREAM) {
        const int per = (int)(gridDim.x >> 3);
        blk = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    }
    if (blk >= nblk) return;  // uniform per workgroup (grid rounded up to 8)
    const int r0 = blk * kSpmmRows;
    const int r = r0 + threadIdx.x;
    const int rend = min(r0 + kSpmmRows, n);
    const int e0 = rowptr[r0];
    const int e1 = rowptr[rend];
    const bool staged = (e1 - e0) <= kSpmmCap;
    if (staged) {
        for (int e = e0 + threadIdx.x; e < e1; e += kSpmmRows) {
            if (STREAM) {
                s_col[e - e0] = __builtin_nontemporal_load(col + e);
                s_val[e - e0] = __builtin_nontemporal_load(val + e);
            } else {
                s_col[e - e0] = col[e];
                s_val[e - e0] = val[e];
            }
        }
    }
    __syncthreads();
    int s = 0, t = 0;
    if (r < n) {
        s = rowptr[r];
        t = rowptr[r + 1];
    }
    const bool active = r < n && t - s <= long_thr;  // long rows: k_spmm_long
    if (!active) t = s;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int rs = r < n ? r : 0;  // a row index that is safe to gather
    for (int c0 = 0; c0 < k; c0 += KC) {
        const T *xb[KC];
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            int c = min(c0 + kk, k - 1);
            int64_t xc = g.use_idx ? (int64_t)g.idx[c] : (int64_t)c;
            xb[kk] = X + xc * ldx;
        }
        T acc[KC];
#pragma unroll
        for (int kk = 0; kk < KC; kk++) acc[kk] = T(0);
        for (int e = s; e < t; e += kSpmmBatch) {
            const int cnt = t - e;  // > 0; the batch holds min(cnt, kSpmmBatch)
            int cj[kSpmmBatch];
            T v[kSpmmBatch];
#pragma unroll
            for (int u = 0; u < kSpmmBatch; u++) {
                if (u < cnt) {
                    if (staged) {
                        cj[u] = s_col[e + u - e0];
                        v[u] = s_val[e + u - e0];
                    } else if (STREAM) {
                        cj[u] = __builtin_nontemporal_load(col + e + u);
                        v[u] = __builtin_nontemporal_load(val + e + u);
                    } else {
                        cj[u] = col[e + u];
                        v[u] = val[e + u];
                    }
                } else {
                    cj[u] = rs;
                    v[u] = T(0);
                }
            }
            T xv[kSpmmBatch][KC];
#pragma unroll
            for (int u = 0; u < kSpmmBatch; u++)
#pragma unroll
                for (int kk = 0; kk < KC; kk++) xv[u][kk] = (u < cnt) ? xb[kk][cj[u]] : T(0);
#pragma unroll
            for (int u = 0; u < kSpmmBatch; u++)
                if (u < cnt) {
#pragma unroll
                    for (int kk = 0; kk < KC; kk++) acc[kk] = acc[kk] + v[u] * xv[u][kk];
                }
        }
        double d[KC];
#pragma unroll
        for (int kk = 0; kk < KC; kk++) {
            const int c = c0 + kk;
            if (active && c < k) {
                if (STREAM) __builtin_nontemporal_store(acc[kk], Y + (int64_t)c * ldy + r);
                else Y[(int64_t)c * ldy + r] = acc[kk];
            }
            if (DOT) d[kk] = (active && c < k) ? (double)xb[kk][r] * (double)acc[kk] : 0.0;
        }
        if (DOT) {
            const double colsum = wave_colsum<KC>(d, lane);
            constexpr int L = KC == 1 ? 0 : KC == 2 ? 1 : KC == 4 ? 2 : KC == 8 ? 3 : KC == 16 ? 4 : 5;
            const int c = c0 + (lane >> (6 - L));
            if ((lane & ((64 >> L) - 1)) == 0 && c < k) sdot[c][wid] = colsum;
        }
    }
    if (DOT) {
        __syncthreads();
        for (int c = threadIdx.x; c < k; c += kSpmmRows)
            partial[(int64_t)c * pstride + blk] = (sdot[c][0] + sdot[c][1]) + (sdot[c][2] + sdot[c][3]);
    }
}

// The rows longer than kLongRow: one wavefront per row (four per workgroup),
// lane l taking column c0 + l, so every column's sum runs over the row in CSR
// order from 0, as in k_spmm (bit-exact); the row's entries go in batches of
// kLongBatch whose column indices and values are read once for the
// wavefront and whose X gathers are all issued before the first product is
// added.  With DOT, the row's X(r,j) Y(r,j) goes to partial[j * pstride +
// pbase + li].
constexpr int kLongBatch = 8;

template <typename T, bool DOT>
__global__ __launch_bounds__(kSpmmRows) void k_spmm_long(const int *__restrict__ rowptr, const int *__restrict__ col,
                                                         const T *__restrict__ val, const T *__restrict__ X,
                                                         int64_t ldx, T *__restrict__ Y, int64_t ldy, int k,
                                                         ColScalars g, const int *__restrict__ long_rows, int nlong,
                                                         double *__restrict__ partial, int pstride, int pbase)
{
    const int lane = threadIdx.x & 63;
    const int li = blockIdx.x * (kSpmmRows / 64) + (threadIdx.x >> 6);
    if (li >= nlong) return;  // uniform per wavefront; no workgroup barrier here
    const int r = long_rows[li];
    const int s = rowptr[r], t = rowptr[r + 1];
    for (int c0 = 0; c0 < k; c0 += 64) {
        const int c = c0 + lane;
        const bool live = c < k;
        const int cc = live ? c : k - 1;
        const T *xc = X + (g.use_idx ? (int64_t)g.idx[cc] : (int64_t)cc) * ldx;
        T acc = T(0);
        for (int e = s; e < t; e += kLongBatch) {
            const int cnt = t - e;
            int cj[kLongBatch];
            T v[kLongBatch], xv[kLongBatch];
#pragma unroll
            for (int u = 0; u < kLongBatch; u++) {
                cj[u] = u < cnt ? col[e + u] : r;
                v[u] = u < cnt ? val[e + u] : T(0);
            }
#pragma unroll
            for (int u = 0; u < kLongBatch; u++) xv[u] = xc[cj[u]];
#pragma unroll
            for (int u = 0; u < kLongBatch; u++)
                if (u < cnt) acc = acc + v[u] * xv[u];
        }
        if (live) {
            Y[(int64_t)c * ldy + r] = acc;
            if (DOT) partial[(int64_t)c * pstride + pbase + li] = (double)xc[r] * (double)acc;
        }
    }
}

// ------------------------------------------------- SpMM, staged row blocks
// The column-gather kernel above reads, for every stored entry and column,
// X(col, j) straight from memory; on a stencil the seven gathers of a row hit
// the same few cache lines, but the L1 keeps almost none of them, so each
// column costs about eight times its own bytes in L2 requests.  Here a row
// block's workgroup first copies every X row its entries read -- its own 256
// rows, then the other rows its entries reference ("extras", listed once per
// block at matrix creation, sorted) -- for KC columns into LDS, with
// coalesced loads, and the rows then read their entries' values from LDS
// through 16-bit local column indices (also built at creation).  Each row's
// first kHaloMaxE entries (local indices and values) stay in registers for all
// columns.  Every sum still runs over the row in CSR order from 0, so Y is
// bit-identical to k_spmm's.  Rows longer than long_thr are left to
// k_spmm_long.  Row blocks are numbered so that each XCD walks a contiguous
// range of rows (a stencil's plane neighbours stay in its L2).
constexpr int kHaloMaxE = 8;

// KC consecutive values of one LDS slot, moved as one vector
template <typename T, int KC>
struct alignas(KC * sizeof(T) >= 16 ? 16 : KC * sizeof(T)) SlotVec {
    T v[KC];
};

template <typename T, bool DOT, int KC>
__global__ __launch_bounds__(kSpmmRows) void k_spmm_halo(int n, const int *__restrict__ rowptr,
                                                         const unsigned short *__restrict__ lcol,
                                                         const T *__restrict__ val, const int *__restrict__ hptr,
                                                         const int *__restrict__ hcol, const T *__restrict__ X,
                                                         int64_t ldx, T *__restrict__ Y, int64_t ldy, int k,
                                                         ColScalars g, double *__restrict__ partial, int nblk,
                                                         int long_thr, int pstride, int hcap)
{
    // s_x holds slot-major KC-column rows: slot j, column cc at s_x[j * KC + cc]
    extern __shared__ __align__(16) unsigned char halo_smem[];
    T *s_x = reinterpret_cast<T *>(halo_smem);                    // [hcap][KC]
    int *s_h = reinterpret_cast<int *>(s_x + (size_t)KC * hcap);  // [hcap - kSpmmRows]
    __shared__ double sdot[DOT ? kMaxCols : 1][4];
    const int per = (int)(gridDim.x >> 3);
    const int blk = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (blk >= nblk) return;  // uniform per workgroup (grid rounded up to 8)
    const int tid = threadIdx.x, r0 = blk * kSpmmRows, r = r0 + tid;
    const int nown = min(kSpmmRows, n - r0);
    const int h0 = hptr[blk], nh = hptr[blk + 1] - h0;
    for (int j = tid; j < nh; j += kSpmmRows) s_h[j] = hcol[h0 + j];
    int s = 0, t = 0;
    if (r < n) {
        s = rowptr[r];
        t = rowptr[r + 1];
    }
    const bool active = r < n && t - s <= long_thr;  // long rows: k_spmm_long
    const int cnt = active ? t - s : 0;
    int lc[kHaloMaxE];
    T v[kHaloMaxE];
#pragma unroll
    for (int u = 0; u < kHaloMaxE; u++) {
        lc[u] = u < cnt ? (int)lcol[s + u] : 0;
        v[u] = u < cnt ? val[s + u] : T(0);
    }
    const int lane = tid & 63, wid = tid >> 6;
    using SV = SlotVec<T, KC>;
    SV *sv = reinterpret_cast<SV *>(s_x);
    __syncthreads();
    // the own row and the first extra slot of this thread for the next KC
    // columns are loaded into registers while the current columns are
    // multiplied, so that a workgroup's loads overlap its own work
    const int row1 = tid < nh ? s_h[tid] : 0;
    SV pown, pext;
    auto xcols = [&](int c0, const T *(&xc)[KC]) {
#pragma unroll
        for (int cc = 0; cc < KC; cc++) {
            const int c = min(c0 + cc, k - 1);
            xc[cc] = X + (g.use_idx ? (int64_t)g.idx[c] : (int64_t)c) * ldx;
        }
    };
    auto prefetch = [&](int c0) {
        const T *xc[KC];
        xcols(c0, xc);
        if (tid < nown) {
#pragma unroll
            for (int cc = 0; cc < KC; cc++) pown.v[cc] = xc[cc][r0 + tid];
        }
        if (tid < nh) {
#pragma unroll
            for (int cc = 0; cc < KC; cc++) pext.v[cc] = xc[cc][row1];
        }
    };
    prefetch(0);
    for (int c0 = 0; c0 < k; c0 += KC) {
        const T *xc[KC];
        xcols(c0, xc);
        if (tid < nown) sv[tid] = pown;
        if (tid < nh) sv[kSpmmRows + tid] = pext;
        // extras beyond one per thread (blocks with more than 256) load here
        for (int j = tid + kSpmmRows; j < nh; j += kSpmmRows) {
            const int row = s_h[j];
            SV w;
#pragma unroll
            for (int cc = 0; cc < KC; cc++) w.v[cc] = xc[cc][row];
            sv[kSpmmRows + j] = w;
        }
        __syncthreads();
        if (c0 + KC < k) prefetch(c0 + KC);
        T acc[KC];
#pragma unroll
        for (int cc = 0; cc < KC; cc++) acc[cc] = T(0);
#pragma unroll
        for (int u = 0; u < kHaloMaxE; u++)
            if (u < cnt) {
                const SV x = sv[lc[u]];
#pragma unroll
                for (int cc = 0; cc < KC; cc++) acc[cc] = acc[cc] + v[u] * x.v[cc];
            }
        for (int e = s + kHaloMaxE; e < s + cnt; e++) {
            const SV x = sv[lcol[e]];
            const T ve = val[e];
#pragma unroll
            for (int cc = 0; cc < KC; cc++) acc[cc] = acc[cc] + ve * x.v[cc];
        }
        double d[KC];
        const SV xo = sv[tid];  // own row (written by this thread when active)
#pragma unroll
        for (int cc = 0; cc < KC; cc++) {
            const int c = c0 + cc;
            if (active && c < k) Y[(int64_t)c * ldy + r] = acc[cc];
            if (DOT) d[cc] = (active && c < k) ? (double)xo.v[cc] * (double)acc[cc] : 0.0;
        }
        if (DOT) {
            const double colsum = wave_colsum<KC>(d, lane);
            constexpr int L = KC == 1 ? 0 : KC == 2 ? 1 : KC == 4 ? 2 : KC == 8 ? 3 : KC == 16 ? 4 : 5;
            const int c = c0 + (lane >> (6 - L));
            if ((lane & ((64 >> L) - 1)) == 0 && c < k) sdot[c][wid] = colsum;
        }
        __syncthreads();  // the next chunk overwrites s_x
    }
    if (DOT) {
        for (int c = tid; c < k; c += kSpmmRows)
            partial[(int64_t)c * pstride + blk] = (sdot[c][0] + sdot[c][1]) + (sdot[c][2] + sdot[c][3]);
    }
}

// ------------------------------------------------- SpMM, row-major gather
// Matrices whose column indices scatter (random graphs) make the column-major
// gather of k_spmm touch one cache line per stored entry and column, and
// matrices whose rows average more than 8 entries overflow its LDS stage, so
// that each thread reads its own row's indices and values.  For both
// (HIP_MAT::gather_rows, chosen at creation) the k columns are first copied into a
// row-major staging block Xt, row r holding X(r, 0..kp) (kp: k rounded up to
// the lane group G), and k_spmm_rows gathers whole row segments: G lanes of a
// wavefront take G consecutive columns of one row, so a stored entry costs
// one contiguous G * sizeof(T) read.  Every sum still runs over its row in
// CSR order from 0, so Y is bit-identical to the column-major path.

// Xt(r, c) = X(:, c)(r) (c < k; 0 for k <= c < kp): 64 x 64 tiles through LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_rows_from_cols(int n, const T *__restrict__ X, int64_t ldx, int k,
                                                        ColScalars g, int kp, T *__restrict__ Xt)
{
    __shared__ T tile[64][65];
    const int r0 = blockIdx.x * 64, cb = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int cc = ty; cc < 64; cc += 4) {
        const int c = cb + cc, r = r0 + tx;
        T v = T(0);
        if (c < k && r < n) v = X[(g.use_idx ? (int64_t)g.idx[c] : (int64_t)c) * ldx + r];
        tile[cc][tx] = v;
    }
    __syncthreads();
    const int kpe = min(64, kp - cb), rows = min(64, n - r0);
    for (int q = threadIdx.x; q < rows * kpe; q += 256) {
        const int rr = q / kpe, cc = q - rr * kpe;
        Xt[(int64_t)(r0 + rr) * kp + cb + cc] = tile[cc][rr];
    }
}

__device__ __forceinline__ float lane_bcast(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

__device__ __forceinline__ double lane_bcast(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

constexpr int kRowsBatch = 8;       // gathers in flight per lane, short rows
constexpr int kLongRowsBatch = 32;  // gathers in flight per lane, long rows

// Sum over the CSR row [s, t) of val(e) * Xt(col(e), cc), in CSR order from
// 0, for a whole wavefront (lane = column): the row's column indices and
// values are read 64 at a time and broadcast from the lane holding them
// (readlane), so the gathers have uniform row addresses and kLongRowsBatch
// of them are in flight.  TA is the accumulation type (T for the SpMM, FP64
// for the residual check); rfill is a valid row for the idle slots.  Called
// by all 64 lanes together.
template <typename TA, typename TV, typename TX>
__device__ __forceinline__ TA wave_row_sum(const int *__restrict__ col, const TV *__restrict__ val, int s, int t,
                                           const TX *__restrict__ Xt, int kp, int cc, int rfill, int lane)
{
    TA acc = TA(0);
    for (int e = s; e < t; e += 64) {
        const int cnt = min(64, t - e);
        const int cv = lane < cnt ? col[e + lane] : rfill;
        const TV vv = lane < cnt ? val[e + lane] : TV(0);
#pragma unroll
        for (int h = 0; h < 64; h += kLongRowsBatch) {
            if (h < cnt) {
                TX xv[kLongRowsBatch];
#pragma unroll
                for (int u = 0; u < kLongRowsBatch; u++) {
                    const int cj = __builtin_amdgcn_readlane(cv, h + u);
                    xv[u] = Xt[(int64_t)cj * kp + cc];
                }
#pragma unroll
                for (int u = 0; u < kLongRowsBatch; u++)
                    if (h + u < cnt) acc = acc + lane_bcast(vv, h + u) * (TA)xv[u];
            }
        }
    }
    return acc;
}

// The same sum for one lane (row) of a lane group: kRowsBatch entries'
// gathers in flight; rfill is a valid row for the idle slots.
template <typename TA, typename TV, typename TX>
__device__ __forceinline__ TA lane_row_sum(const int *__restrict__ col, const TV *__restrict__ val, int s, int t,
                                           const TX *__restrict__ Xt, int kp, int c, int rfill)
{
    TA acc = TA(0);
    for (int e = s; e < t; e += kRowsBatch) {
        const int cnt = t - e;
        int cj[kRowsBatch];
        TV v[kRowsBatch];
        TX xv[kRowsBatch];
#pragma unroll
        for (int u = 0; u < kRowsBatch; u++) {
            cj[u] = u < cnt ? col[e + u] : rfill;
            v[u] = u < cnt ? val[e + u] : TV(0);
        }
#pragma unroll
        for (int u = 0; u < kRowsBatch; u++) xv[u] = Xt[(int64_t)cj[u] * kp + c];
#pragma unroll
        for (int u = 0; u < kRowsBatch; u++)
            if (u < cnt) acc = acc + v[u] * (TA)xv[u];
    }
    return acc;
}

// Rows handled by one workgroup of k_spmm_rows<G>: four wavefronts of
// max(16, 64 / G) rows each.
__host__ __device__ constexpr int rows_per_wg(int G) { return 4 * (64 / G > 16 ? 64 / G : 16); }

// Y(:, c0 + 0..G) = A X from the staging block Xt (grid.y: chunks of 64
// columns when G = 64).  Workgroups below nlb take the rows longer than
// long_thr, one wavefront per row with lane l on column c0 + l; the row's
// column indices and values are read 64 at a time and broadcast from the
// lane holding them (readlane), so the gathers have uniform row addresses
// and kLongRowsBatch of them are in flight.  The other workgroups take
// rows_per_wg(G) consecutive rows: lane l works on column c0 + l % G of row
// l / G (64 / G rows in flight per wavefront), the sums land in LDS and
// leave as contiguous column segments.  With DOT, sum_r X(r,c) Y(r,c) goes
// to partial[c * pstride + blk] per row workgroup and to
// partial[c * pstride + nblk + li] per long row.
template <typename T, bool DOT, int G>
__global__ __launch_bounds__(256) void k_spmm_rows(int n, const int *__restrict__ rowptr,
                                                   const int *__restrict__ col, const T *__restrict__ val,
                                                   const T *__restrict__ Xt, int kp, T *__restrict__ Y,
                                                   int64_t ldy, int k, const int *__restrict__ long_rows,
                                                   int nlong, int nlb, int long_thr,
                                                   double *__restrict__ partial, int pstride, int nblk)
{
    static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0, "G is a power of two");
    constexpr int SG = 64 / G;
    constexpr int RWG = rows_per_wg(G);
    constexpr int RPW = RWG / 4;
    __shared__ T sY[G][RWG + 1];
    __shared__ unsigned char skip[RWG];
    __shared__ double sdot[DOT ? 4 : 1][G];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c0 = blockIdx.y * 64;
    if ((int)blockIdx.x < nlb) {
        const int li = blockIdx.x * 4 + wid;
        if (li >= nlong) return;  // uniform per wavefront; this branch has no barrier
        const int r = long_rows[li];
        const int s = rowptr[r], t = rowptr[r + 1];
        const int c = c0 + lane;
        const bool live = c < k;
        const int cc = min(c, kp - 1);
        const T acc = wave_row_sum<T, T, T>(col, val, s, t, Xt, kp, cc, r, lane);
        if (live) {
            Y[(int64_t)c * ldy + r] = acc;
            if (DOT) partial[(int64_t)c * pstride + nblk + li] = (double)Xt[(int64_t)r * kp + cc] * (double)acc;
        }
        return;
    }
    const int blk = blockIdx.x - nlb;
    const int r0 = blk * RWG;
    const int cl = lane % G, sg = lane / G;
    const int c = c0 + cl;  // < kp: kp = G, or a multiple of 64 when G = 64
    const bool live = c < k;
    double dsum = 0.0;
    for (int i = sg; i < RPW; i += SG) {
        const int rl = wid * RPW + i, r = r0 + rl;
        int s = 0, t = 0;
        if (r < n) {
            s = rowptr[r];
            t = rowptr[r + 1];
        }
        const bool act = r < n && t - s <= long_thr;
        if (!act) t = s;
        const int rs = r < n ? r : 0;
        const T acc = lane_row_sum<T, T, T>(col, val, s, t, Xt, kp, c, rs);
        if (cl == 0) skip[rl] = act ? 0 : 1;
        sY[cl][rl] = acc;
        if (DOT && act && live) dsum += (double)Xt[(int64_t)r * kp + c] * (double)acc;
    }
    if (DOT) {
#pragma unroll
        for (int m = G; m < 64; m <<= 1) dsum += __shfl_xor(dsum, m, 64);
        if (sg == 0) sdot[wid][cl] = dsum;
    }
    __syncthreads();
    const int ncol = min(G, k - c0);
    for (int q = threadIdx.x; q < ncol * RWG; q += 256) {
        const int cc = q / RWG, rr = q - cc * RWG;
        const int r = r0 + rr;
        if (r < n && !skip[rr]) Y[(int64_t)(c0 + cc) * ldy + r] = sY[cc][rr];
    }
    if (DOT && (int)threadIdx.x < ncol)
        partial[(int64_t)(c0 + threadIdx.x) * pstride + blk] =
            (sdot[0][threadIdx.x] + sdot[1][threadIdx.x]) + (sdot[2][threadIdx.x] + sdot[3][threadIdx.x]);
}

// ---------------------------------------------------------------- residual
// r = A x - lambda B x (x when B is absent), in FP64 arithmetic whatever the
// storage precision of x, rounded as VecAxpby(-lambda, Bx, 1.0, Ax)
// (gcge_eigsol.c:579-588); partial sums of r.r, x.x and Bx.Bx per column.
template <typename TX, bool HASB>
__global__ __launch_bounds__(kSpmmRows) void k_residual(int n, const int *__restrict__ arp,
                                                        const int *__restrict__ acol,
                                                        const double *__restrict__ aval,
                                                        const int *__restrict__ brp,
                                                        const int *__restrict__ bcol,
                                                        const double *__restrict__ bval,
                                                        const TX *__restrict__ X, int64_t ldx, int k,
                                                        ColScalars g, double *__restrict__ partial)
{
    // The workgroup's rows of A are staged in LDS once and serve every
    // column; columns go in chunks of kResKC, each thread one row, and the
    // three sums of every column are reduced per wavefront into LDS and
    // combined after the last chunk (same order as block_sum).
    constexpr int kResKC = 4;
    __shared__ int s_col[kSpmmCap];
    __shared__ double s_val[kSpmmCap];
    __shared__ double sred[3 * kMaxCols][4];
    const int blk = blockIdx.x;
    const int r0 = blk * kSpmmRows;
    const int r = r0 + threadIdx.x;
    const int rend = min(r0 + kSpmmRows, n);
    const int e0 = arp[r0];
    const int e1 = arp[rend];
    const bool staged = (e1 - e0) <= kSpmmCap;
    if (staged) {
        for (int e = e0 + threadIdx.x; e < e1; e += kSpmmRows) {
            s_col[e - e0] = acol[e];
            s_val[e - e0] = aval[e];
        }
    }
    __syncthreads();
    const bool active = r < n;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int s = 0, t = 0;
    if (active) {
        s = arp[r];
        t = arp[r + 1];
    }
    for (int c0 = 0; c0 < k; c0 += kResKC) {
        const TX *xj[kResKC];
#pragma unroll
        for (int kk = 0; kk < kResKC; kk++) xj[kk] = X + (int64_t)g.idx[min(c0 + kk, k - 1)] * ldx;
        double ax[kResKC];
#pragma unroll
        for (int kk = 0; kk < kResKC; kk++) ax[kk] = 0.0;
        if (active) {
            if (staged) {
                for (int e = s; e < t; e++) {
                    const int cj = s_col[e - e0];
                    const double v = s_val[e - e0];
#pragma unroll
                    for (int kk = 0; kk < kResKC; kk++) ax[kk] = ax[kk] + v * (double)xj[kk][cj];
                }
            } else {
                for (int e = s; e < t; e++) {
                    const int cj = acol[e];
                    const double v = aval[e];
#pragma unroll
                    for (int kk = 0; kk < kResKC; kk++) ax[kk] = ax[kk] + v * (double)xj[kk][cj];
                }
            }
        }
#pragma unroll
        for (int kk = 0; kk < kResKC; kk++) {
            const int c = c0 + kk;
            if (c >= k) break;
            double rr = 0.0, xx = 0.0, bb = 0.0;
            if (active) {
                const double xr = (double)xj[kk][r];
                double bx;
                if (HASB) {
                    bx = 0.0;
                    for (int e = brp[r]; e < brp[r + 1]; e++) bx = bx + bval[e] * (double)xj[kk][bcol[e]];
                } else {
                    bx = xr;
                }
                const double res = axpby_el<double>(-g.a[c], bx, 1.0, ax[kk]);
                rr = res * res;
                xx = xr * xr;
                bb = bx * bx;
            }
            rr = wave_sum(rr);
            xx = wave_sum(xx);
            bb = wave_sum(bb);
            if (lane == 0) {
                sred[3 * c + 0][wid] = rr;
                sred[3 * c + 1][wid] = xx;
                sred[3 * c + 2][wid] = bb;
            }
        }
    }
    __syncthreads();
    const int nb = gridDim.x;
    for (int q = threadIdx.x; q < 3 * k; q += kSpmmRows)
        partial[(int64_t)q * nb + blk] = (sred[q][0] + sred[q][1]) + (sred[q][2] + sred[q][3]);
}

// r = A x - lambda B x (Bx = x when B is absent) on the row-major gather
// path: the FP64 row sums of k_residual (CSR order from 0, x converted to
// FP64 per entry) gathered from the staged rows Xt, in the workgroup layout
// of k_spmm_rows (A's long rows one wavefront each; B has no row longer than
// long_thr).  The sums r.r, x.x and Bx.Bx of column c go to
// partial[(3c + q) * pstride + blk] per row workgroup and to
// partial[(3c + q) * pstride + nblk + li] per long row.
template <typename TX, int G, bool HASB>
__global__ __launch_bounds__(256) void k_residual_rows(int n, const int *__restrict__ rowptr,
                                                       const int *__restrict__ col, const double *__restrict__ val,
                                                       const int *__restrict__ brp, const int *__restrict__ bcol,
                                                       const double *__restrict__ bval, const TX *__restrict__ Xt,
                                                       int kp, int k, ColScalars g,
                                                       const int *__restrict__ long_rows, int nlong, int nlb,
                                                       int long_thr, double *__restrict__ partial, int pstride,
                                                       int nblk)
{
    constexpr int SG = 64 / G;
    constexpr int RWG = rows_per_wg(G);
    constexpr int RPW = RWG / 4;
    __shared__ double sred[4][G][3];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c0 = blockIdx.y * 64;
    if ((int)blockIdx.x < nlb) {
        const int li = blockIdx.x * 4 + wid;
        if (li >= nlong) return;  // uniform per wavefront; this branch has no barrier
        const int r = long_rows[li];
        const int c = c0 + lane;
        const int cc = min(c, kp - 1);
        const double ax = wave_row_sum<double, double, TX>(col, val, rowptr[r], rowptr[r + 1], Xt, kp, cc, r, lane);
        const double xr = (double)Xt[(int64_t)r * kp + cc];
        const double bx =
            HASB ? wave_row_sum<double, double, TX>(bcol, bval, brp[r], brp[r + 1], Xt, kp, cc, r, lane) : xr;
        if (c < k) {
            const double res = axpby_el<double>(-g.a[c], bx, 1.0, ax);
            partial[(int64_t)(3 * c + 0) * pstride + nblk + li] = res * res;
            partial[(int64_t)(3 * c + 1) * pstride + nblk + li] = xr * xr;
            partial[(int64_t)(3 * c + 2) * pstride + nblk + li] = bx * bx;
        }
        return;
    }
    const int blk = blockIdx.x - nlb;
    const int r0 = blk * RWG;
    const int cl = lane % G, sg = lane / G;
    const int c = c0 + cl;  // < kp
    const bool live = c < k;
    double srr = 0.0, sxx = 0.0, sbb = 0.0;
    for (int i = sg; i < RPW; i += SG) {
        const int r = r0 + wid * RPW + i;
        int s = 0, t = 0, bs = 0, bt = 0;
        if (r < n) {
            s = rowptr[r];
            t = rowptr[r + 1];
            if (HASB) {
                bs = brp[r];
                bt = brp[r + 1];
            }
        }
        const bool act = r < n && t - s <= long_thr;
        if (!act) t = s, bt = bs;
        const int rs = r < n ? r : 0;
        const double ax = lane_row_sum<double, double, TX>(col, val, s, t, Xt, kp, c, rs);
        const double bx0 = HASB ? lane_row_sum<double, double, TX>(bcol, bval, bs, bt, Xt, kp, c, rs) : 0.0;
        if (act && live) {
            const double xr = (double)Xt[(int64_t)r * kp + c];
            const double bx = HASB ? bx0 : xr;
            const double res = axpby_el<double>(-g.a[c], bx, 1.0, ax);
            srr += res * res;
            sxx += xr * xr;
            sbb += bx * bx;
        }
    }
#pragma unroll
    for (int m = G; m < 64; m <<= 1) {
        srr += __shfl_xor(srr, m, 64);
        sxx += __shfl_xor(sxx, m, 64);
        sbb += __shfl_xor(sbb, m, 64);
    }
    if (sg == 0) {
        sred[wid][cl][0] = srr;
        sred[wid][cl][1] = sxx;
        sred[wid][cl][2] = sbb;
    }
    __syncthreads();
    const int q = threadIdx.x;
    if (q < min(G, k - c0)) {
        const int cq = c0 + q;
#pragma unroll
        for (int j = 0; j < 3; j++)
            partial[(int64_t)(3 * cq + j) * pstride + blk] =
                (sred[0][q][j] + sred[1][q][j]) + (sred[2][q][j] + sred[3][q][j]);
    }
}

// ---------------------------------------------------------------- CG update
// x += alpha p (VecAxpby(alpha, p, 1.0, x)); r -= alpha w
// (VecAxpby(-alpha, w, 1.0, r)); partial r.r  (gcge_cg.c:250-276)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(int64_t n, T *__restrict__ x, int64_t ldx,
                                                      const T *__restrict__ p, int64_t ldp,
                                                      const T *__restrict__ w, int64_t ldw,
                                                      T *__restrict__ r, int64_t ldr, ColScalars s,
                                                      double *__restrict__ partial)
{
    __shared__ double sh[4];
    const int j = blockIdx.y;
    const T al = (T)s.a[j];
    T *xj = x + (int64_t)j * ldx;
    const T *pj = p + (int64_t)j * ldp;
    const T *wj = w + (int64_t)j * ldw;
    T *rj = r + (int64_t)j * ldr;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        xj[i] = axpby_el<T>(al, pj[i], T(1), xj[i]);
        const T rn = axpby_el<T>(-al, wj[i], T(1), rj[i]);
        rj[i] = rn;
        acc = fma((double)rn, (double)rn, acc);
    }
    const double bs = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = bs;
}

// ---------------------------------------------------------------- split CG steps
// r -= alpha w (VecAxpby(-alpha, w, 1.0, r)); partial r.r  (gcge_cg.c:250-276)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_resid(int64_t n, const T *__restrict__ w, int64_t ldw,
                                                     T *__restrict__ r, int64_t ldr, ColScalars s,
                                                     double *__restrict__ partial)
{
    __shared__ double sh[4];
    const int j = blockIdx.y;
    const T al = (T)s.a[j];
    const T *wj = w + (int64_t)j * ldw;
    T *rj = r + (int64_t)j * ldr;
    double acc = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const T rn = axpby_el<T>(-al, wj[i], T(1), rj[i]);
        rj[i] = rn;
        acc = fma((double)rn, (double)rn, acc);
    }
    const double bs = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[(int64_t)j * gridDim.x + blockIdx.x] = bs;
}

// Jacobi-preconditioned residual step: r -= alpha w (same rounding as
// k_cg_resid), partial r.z (row 2j) and r.r (row 2j + 1) with z = dinv .* r
// rounded as k_cg_search<T, true> forms it
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_resid(int64_t n, const T *__restrict__ w, int64_t ldw,
                                                      T *__restrict__ r, int64_t ldr, const T *__restrict__ dinv,
                                                      ColScalars s, double *__restrict__ partial)
{
    __shared__ double sh[4];
    const int j = blockIdx.y;
    const T al = (T)s.a[j];
    const T *wj = w + (int64_t)j * ldw;
    T *rj = r + (int64_t)j * ldr;
    double arz = 0.0, arr = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
        const T rn = axpby_el<T>(-al, wj[i], T(1), rj[i]);
        const T zn = rn * dinv[i];
        rj[i] = rn;
        arz = fma((double)rn, (double)zn, arz);
        arr = fma((double)rn, (double)rn, arr);
    }
    const double brz = block_sum(arz, sh);
    const double brr = block_sum(arr, sh);
    if (threadIdx.x == 0) {
        partial[(int64_t)(2 * j) * gridDim.x + blockIdx.x] = brz;
        partial[(int64_t)(2 * j + 1) * gridDim.x + blockIdx.x] = brr;
    }
}

// dinv[i] = 1 / a_ii; rows without a nonzero diagonal entry get 1 (no scaling)
__global__ __launch_bounds__(kBlock) void k_diag_inv(int n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ col, const double *__restrict__ val,
                                                     double *__restrict__ d64, float *__restrict__ d32)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int e = rowptr[i]; e < rowptr[i + 1]; e++)
        if (col[e] == i) d = d + val[e];
    const double inv = (d != 0.0) ? 1.0 / d : 1.0;
    d64[i] = inv;
    d32[i] = (float)inv;
}

// x += alpha p (VecAxpby(alpha, p, 1.0, x)) for the columns j < kx; then
// p = r + beta p (MultiVecAxpbyDiag(1, r, beta, p)) for the columns j < kp
// With PRE, the search direction is built from z = dinv .* r (the Jacobi-
// preconditioned residual, formed here rather than stored).
template <typename T, bool PRE>
__global__ __launch_bounds__(kBlock) void k_cg_search(int64_t n, T *__restrict__ x, int64_t ldx,
                                                      T *__restrict__ p, int64_t ldp, const T *__restrict__ r,
                                                      int64_t ldr, int kx, int kp, ColScalars s,
                                                      const T *__restrict__ dinv)
{
    const int j = blockIdx.y;
    const T al = (T)s.a[j], be = (T)s.b[j];
    T *xj = x + (int64_t)j * ldx;
    T *pj = p + (int64_t)j * ldp;
    const T *rj = r + (int64_t)j * ldr;
    const bool dx = j < kx, dp = j < kp;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    if (dx && dp) {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
            const T pv = pj[i];
            xj[i] = axpby_el<T>(al, pv, T(1), xj[i]);
            pj[i] = axpby_el<T>(T(1), PRE ? rj[i] * dinv[i] : rj[i], be, pv);
        }
    } else if (dx) {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
            xj[i] = axpby_el<T>(al, pj[i], T(1), xj[i]);
    } else if (dp) {
        for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
            pj[i] = axpby_el<T>(T(1), PRE ? rj[i] * dinv[i] : rj[i], be, pj[i]);
    }
}

// ---------------------------------------------------------------- gradient
// g_e = sum_i w_i x_i[row_e] x_i[col_e], one thread per stored entry
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sampled_outer(int nnz, const int *__restrict__ rowidx,
                                                          const int *__restrict__ col,
                                                          const T *__restrict__ X, int64_t ldx,
                                                          int k, ColScalars s, double *__restrict__ g)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= nnz) return;
    const int r = rowidx[e], c = col[e];
    double acc = 0.0;
    for (int i = 0; i < k; i++) {
        const T *xi = X + (int64_t)i * ldx;
        acc = fma(s.a[i], (double)xi[r] * (double)xi[c], acc);
    }
    g[e] += acc;
}

__global__ __launch_bounds__(kBlock) void k_csr_gather_sum(int nparam, const int *__restrict__ jrp,
                                                           const int *__restrict__ jnz,
                                                           const double *__restrict__ jval,
                                                           const double *__restrict__ g,
                                                           double *__restrict__ grad, double scale, int accumulate)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= nparam) return;
    double acc = 0.0;
    for (int t = jrp[p]; t < jrp[p + 1]; t++) acc = fma(jval[t], g[jnz[t]], acc);
    grad[p] = accumulate ? fma(scale, acc, grad[p]) : scale * acc;
}

__global__ __launch_bounds__(kBlock) void k_rowidx(int n, const int *__restrict__ rowptr,
                                                   int *__restrict__ rowidx)
{
    const int r = blockIdx.x * kBlock + threadIdx.x;
    if (r >= n) return;
    for (int e = rowptr[r]; e < rowptr[r + 1]; e++) rowidx[e] = r;
}

__global__ __launch_bounds__(kBlock) void k_to_f32(int64_t n, const double *__restrict__ a,
                                                   float *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

}  // namespace

// ======================================================================
// launchers
// ======================================================================

static ColScalars make_scalars(int k, const double *a, const double *b, const int *idx)
{
    ColScalars s;
    std::memset(&s, 0, sizeof(s));
    for (int j = 0; j < k; j++) {
        s.a[j] = a ? a[j] : 0.0;
        s.b[j] = b ? b[j] : 0.0;
        s.idx[j] = idx ? idx[j] : j;
    }
    s.use_idx = idx ? 1 : 0;
    return s;
}

void launch_axpby(int prec, int64_t n, const void *x, int64_t ldx, void *y, int64_t ldy, int k,
                  const double *a, const double *b, const int *xidx)
{
    if (k <= 0 || n <= 0 || failed()) return;
    ColScalars s = make_scalars(k, a, b, xidx);
    dim3 grid(grid_1d(n, 1024), k);
    double bytes = 0.0;
    for (int j = 0; j < k; j++) bytes += (double)n * elem_size(prec) * ((s.a[j] == 0.0 || s.b[j] == 0.0) ? 2 : 3);
    ProfScope ps(K_AXPBY, bytes, 0.0);
    if (prec == GCGE_HIP_FP64)
        k_axpby<double><<<grid, kBlock, 0, ctx().stream>>>(n, (const double *)x, ldx, (double *)y, ldy, s);
    else
        k_axpby<float><<<grid, kBlock, 0, ctx().stream>>>(n, (const float *)x, ldx, (float *)y, ldy, s);
    check(hipGetLastError(), "k_axpby");
}

static bool reduce_to_host(double *partial, int rows, int nb, double *out)
{
    double *dres = partial + (int64_t)rows * nb;
    k_reduce_rows<<<rows, kBlock, 0, ctx().stream>>>(partial, nb, dres);
    if (!check(hipGetLastError(), "k_reduce_rows")) return false;
    double *h = host_buf(rows);
    if (!h) return false;
    if (!check(hipMemcpyAsync(h, dres, sizeof(double) * rows, hipMemcpyDeviceToHost, ctx().stream),
               "hipMemcpyAsync D2H"))
        return false;
    if (!sync()) return false;
    std::memcpy(out, h, sizeof(double) * rows);
    return true;
}

int dots_blocks(int64_t n) { return grid_1d(n, 256); }

static void launch_dots(int prec, int64_t n, const void *x, int64_t ldx, const void *y, int64_t ldy, int k,
                        double *partial, int nb)
{
    dim3 grid(nb, k);
    if (prec == GCGE_HIP_FP64)
        k_dots<double><<<grid, kBlock, 0, ctx().stream>>>(n, (const double *)x, ldx, (const double *)y, ldy, partial);
    else
        k_dots<float><<<grid, kBlock, 0, ctx().stream>>>(n, (const float *)x, ldx, (const float *)y, ldy, partial);
}

bool dots(int prec, int64_t n, const void *x, int64_t ldx, const void *y, int64_t ldy, int k,
          double *out)
{
    if (k <= 0) return true;
    if (failed()) return false;
    const int nb = dots_blocks(n);
    double *partial = (double *)scratch(sizeof(double) * ((size_t)k * nb + k));
    if (!partial) return false;
    {
        ProfScope ps(K_DOTS, 2.0 * k * (double)n * elem_size(prec), 2.0 * k * (double)n);
        launch_dots(prec, n, x, ldx, y, ldy, k, partial, nb);
    }
    if (!check(hipGetLastError(), "k_dots")) return false;
    return reduce_to_host(partial, k, nb, out);
}

// The same dots with the partial sums in a caller-provided device buffer of
// at least k * dots_blocks(n) + k doubles and the sums copied on the backend
// stream into host memory out (pinned or not); no allocation, no profiler.
bool dots_with(int prec, int64_t n, const void *x, int64_t ldx, const void *y, int64_t ldy, int k, double *partial,
               double *out)
{
    if (k <= 0) return true;
    if (failed()) return false;
    const int nb = dots_blocks(n);
    launch_dots(prec, n, x, ldx, y, ldy, k, partial, nb);
    if (!check(hipGetLastError(), "k_dots")) return false;
    double *dres = partial + (int64_t)k * nb;
    k_reduce_rows<<<k, kBlock, 0, ctx().stream>>>(partial, nb, dres);
    if (!check(hipGetLastError(), "k_reduce_rows")) return false;
    if (!check(hipMemcpyAsync(out, dres, sizeof(double) * k, hipMemcpyDeviceToHost, ctx().stream), "D2H dots"))
        return false;
    return check(hipStreamSynchronize(ctx().stream), "hipStreamSynchronize");
}

void launch_copy(int xprec, const void *x, int64_t ldx, int yprec, void *y, int64_t ldy, int64_t n, int k)
{
    if (k <= 0 || n <= 0 || failed()) return;
    dim3 grid(grid_1d(n, 1024), k);
    hipStream_t st = ctx().stream;
    ProfScope ps(K_COPY, (double)k * n * (elem_size(xprec) + elem_size(yprec)), 0.0);
    if (xprec == GCGE_HIP_FP64 && yprec == GCGE_HIP_FP64)
        k_copy<double, double><<<grid, kBlock, 0, st>>>(n, (const double *)x, ldx, (double *)y, ldy);
    else if (xprec == GCGE_HIP_FP64)
        k_copy<double, float><<<grid, kBlock, 0, st>>>(n, (const double *)x, ldx, (float *)y, ldy);
    else if (yprec == GCGE_HIP_FP64)
        k_copy<float, double><<<grid, kBlock, 0, st>>>(n, (const float *)x, ldx, (double *)y, ldy);
    else
        k_copy<float, float><<<grid, kBlock, 0, st>>>(n, (const float *)x, ldx, (float *)y, ldy);
    check(hipGetLastError(), "k_copy");
}

void launch_swap(int prec, void *x, int64_t ldx, void *y, int64_t ldy, int64_t n, int k)
{
    if (k <= 0 || n <= 0 || failed()) return;
    dim3 grid(grid_1d(n, 1024), k);
    ProfScope ps(K_SWAP, 4.0 * k * (double)n * elem_size(prec), 0.0);
    if (prec == GCGE_HIP_FP64)
        k_swap<double><<<grid, kBlock, 0, ctx().stream>>>(n, (double *)x, ldx, (double *)y, ldy);
    else
        k_swap<float><<<grid, kBlock, 0, ctx().stream>>>(n, (float *)x, ldx, (float *)y, ldy);
    check(hipGetLastError(), "k_swap");
}

void launch_zero_rows(int prec, void *x, int64_t ldx, int64_t n, int k, const unsigned char *mask)
{
    if (k <= 0 || n <= 0 || failed()) return;
    dim3 grid(grid_1d(n, 1024), k);
    if (prec == GCGE_HIP_FP64)
        k_zero_rows<double><<<grid, kBlock, 0, ctx().stream>>>(n, (double *)x, ldx, mask);
    else
        k_zero_rows<float><<<grid, kBlock, 0, ctx().stream>>>(n, (float *)x, ldx, mask);
    check(hipGetLastError(), "k_zero_rows");
}

template <typename T>
static const T *mat_vals(HIP_MAT *A);
template <>
const double *mat_vals<double>(HIP_MAT *A) { return A->val64; }
template <>
const float *mat_vals<float>(HIP_MAT *A) { return mat_val32(A); }

// Row-major gather path (see k_spmm_rows) for kc <= kMaxCols columns.
template <typename T, bool DOT, int G>
static bool spmm_rows_g(HIP_MAT *A, const T *val, const T *Xc, int64_t ldx, ColScalars g, T *Yc, int64_t ldy,
                        int kc, double *out)
{
    const int n = A->n, kp = G < 64 ? G : ((kc + 63) / 64) * 64;
    T *Xt = (T *)row_stage(sizeof(T) * (size_t)n * kp);
    if (!Xt) return false;
    const int RWG = rows_per_wg(G);
    const int nblk = (n + RWG - 1) / RWG, nlong = A->nlong, nlb = (nlong + 3) / 4;
    const int pstride = nblk + (DOT ? nlong : 0);
    double *partial = nullptr;
    if (DOT) {
        partial = (double *)scratch(sizeof(double) * ((size_t)kc * pstride + kc));
        if (!partial) return false;
    }
    {
        const double bytes = (sizeof(T) + 4.0) * A->nnz + 4.0 * (n + 1) + 2.0 * kc * (double)n * sizeof(T);
        const bool f64 = sizeof(T) == 8;
        ProfScope ps(DOT ? (f64 ? K_SPMM_DOT_F64 : K_SPMM_DOT_F32) : (f64 ? K_SPMM_F64 : K_SPMM_F32), bytes,
                     2.0 * kc * (double)A->nnz);
        k_rows_from_cols<T><<<dim3((n + 63) / 64, (kp + 63) / 64), 256, 0, ctx().stream>>>(n, Xc, ldx, kc, g, kp,
                                                                                            Xt);
        k_spmm_rows<T, DOT, G><<<dim3(nlb + nblk, G < 64 ? 1 : kp / 64), 256, 0, ctx().stream>>>(
            n, A->rowptr, A->col, val, Xt, kp, Yc, ldy, kc, A->long_rows, nlong, nlb,
            nlong > 0 ? kLongRow : INT_MAX, partial, pstride, nblk);
    }
    if (!check(hipGetLastError(), "k_spmm_rows")) return false;
    return !DOT || reduce_to_host(partial, kc, pstride, out);
}

template <typename T, bool DOT>
static bool spmm_rows(HIP_MAT *A, const T *val, const T *Xc, int64_t ldx, ColScalars g, T *Yc, int64_t ldy,
                      int kc, double *out)
{
    if (kc > 32) return spmm_rows_g<T, DOT, 64>(A, val, Xc, ldx, g, Yc, ldy, kc, out);
    if (kc > 16) return spmm_rows_g<T, DOT, 32>(A, val, Xc, ldx, g, Yc, ldy, kc, out);
    if (kc > 8) return spmm_rows_g<T, DOT, 16>(A, val, Xc, ldx, g, Yc, ldy, kc, out);
    if (kc > 4) return spmm_rows_g<T, DOT, 8>(A, val, Xc, ldx, g, Yc, ldy, kc, out);
    if (kc > 2) return spmm_rows_g<T, DOT, 4>(A, val, Xc, ldx, g, Yc, ldy, kc, out);
    return spmm_rows_g<T, DOT, 2>(A, val, Xc, ldx, g, Yc, ldy, kc, out);
}

// Columns per LDS pass of k_spmm_halo: 4, or GCGE_HIP_HALO_KC (1, 2, 4 or
// 8), halved while the staged rows exceed the byte budget
// (GCGE_HIP_HALO_LDS, default 40 KiB).
static int halo_kc(const HIP_MAT *A, size_t elem)
{
    const char *e = getenv("GCGE_HIP_HALO_KC");
    const int env_kc = e ? atoi(e) : 0;
    const char *l = getenv("GCGE_HIP_HALO_LDS");
    const long env_lds = l ? atol(l) : 40960;
    // default: 4 columns per pass in both precisions (measured on 3-D and
    // 2-D grids in tile order: 8 FP32 columns were 5-10 % slower, 2 were
    // 10-25 % slower), halved while over the budget
    int kc = (env_kc == 1 || env_kc == 2 || env_kc == 4 || env_kc == 8) ? env_kc : 4;
    if (env_kc != kc)
        while (kc > 1 && (size_t)kc * A->halo_cap * elem > (size_t)env_lds) kc /= 2;
    // never beyond 56 KiB of dynamic LDS (the kernel adds 4 KiB of its own)
    while (kc > 1 && (size_t)kc * A->halo_cap * elem + (size_t)A->halo_cap * 4 > (size_t)56 * 1024) kc /= 2;
    return kc;
}

template <typename T, bool DOT, int KC>
static void spmm_halo_kc(HIP_MAT *A, const T *val, const T *Xc, int64_t ldx, ColScalars g, T *Yc, int64_t ldy,
                         int kc, double *partial, int nblk, int long_thr, int pstride)
{
    const size_t lds = (size_t)KC * A->halo_cap * sizeof(T) + (size_t)(A->halo_cap - kSpmmRows) * sizeof(int);
    k_spmm_halo<T, DOT, KC><<<((nblk + 7) / 8) * 8, kSpmmRows, lds, ctx().stream>>>(
        A->n, A->rowptr, A->lcol, val, A->halo_ptr, A->halo_col, Xc, ldx, Yc, ldy, kc, g, partial, nblk, long_thr,
        pstride, A->halo_cap);
}

template <typename T, bool DOT>
static bool spmm_halo(HIP_MAT *A, const T *val, const T *Xc, int64_t ldx, ColScalars g, T *Yc, int64_t ldy,
                      int kc, double *out)
{
    const int nblk = (A->n + kSpmmRows - 1) / kSpmmRows;
    const int nlong = A->nlong, long_thr = nlong > 0 ? kLongRow : INT_MAX;
    const int pstride = nblk + (DOT ? nlong : 0);
    double *partial = nullptr;
    if (DOT) {
        partial = (double *)scratch(sizeof(double) * ((size_t)kc * pstride + kc));
        if (!partial) return false;
    }
    {
        const double bytes = (sizeof(T) + 4.0) * A->nnz + 4.0 * (A->n + 1) + 2.0 * kc * (double)A->n * sizeof(T);
        const bool f64 = sizeof(T) == 8;
        ProfScope ps(DOT ? (f64 ? K_SPMM_DOT_F64 : K_SPMM_DOT_F32) : (f64 ? K_SPMM_F64 : K_SPMM_F32), bytes,
                     2.0 * kc * (double)A->nnz);
        switch (halo_kc(A, sizeof(T))) {
        case 8: spmm_halo_kc<T, DOT, 8>(A, val, Xc, ldx, g, Yc, ldy, kc, partial, nblk, long_thr, pstride); break;
        case 4: spmm_halo_kc<T, DOT, 4>(A, val, Xc, ldx, g, Yc, ldy, kc, partial, nblk, long_thr, pstride); break;
        case 2: spmm_halo_kc<T, DOT, 2>(A, val, Xc, ldx, g, Yc, ldy, kc, partial, nblk, long_thr, pstride); break;
        default: spmm_halo_kc<T, DOT, 1>(A, val, Xc, ldx, g, Yc, ldy, kc, partial, nblk, long_thr, pstride); break;
        }
        if (nlong > 0)
            k_spmm_long<T, DOT><<<(nlong + kSpmmRows / 64 - 1) / (kSpmmRows / 64), kSpmmRows, 0, ctx().stream>>>(
                A->rowptr, A->col, val, Xc, ldx, Yc, ldy, kc, g, A->long_rows, nlong, partial, pstride, nblk);
    }
    if (!check(hipGetLastError(), "k_spmm_halo")) return false;
    return !DOT || reduce_to_host(partial, kc, pstride, out);
}

template <typename T, bool DOT>
static bool spmm_t(HIP_MAT *A, const void *X, int64_t ldx, void *Y, int64_t ldy, int k, const int *xidx,
                   double *out)
{
    const int nblk = (A->n + kSpmmRows - 1) / kSpmmRows;
    const T *val = mat_vals<T>(A);
    if (!val) return false;
    for (int c0 = 0; c0 < k; c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        ColScalars g = make_scalars(kc, nullptr, nullptr, xidx ? xidx + c0 : nullptr);
        const T *Xc = (const T *)X + (xidx ? 0 : (int64_t)c0 * ldx);
        T *Yc = (T *)Y + (int64_t)c0 * ldy;
        if (A->gather_rows && kc >= 2) {  // one column gathers a line per entry either way
            if (!spmm_rows<T, DOT>(A, val, Xc, ldx, g, Yc, ldy, kc, DOT ? out + c0 : nullptr)) return false;
            continue;
        }
        if (A->halo) {
            if (!spmm_halo<T, DOT>(A, val, Xc, ldx, g, Yc, ldy, kc, DOT ? out + c0 : nullptr)) return false;
            continue;
        }
        // long rows (power-law graphs) go to k_spmm_long; their dot partials
        // follow the workgroups' in each column
        const int nlong = A->nlong, long_thr = nlong > 0 ? kLongRow : INT_MAX;
        const int pstride = nblk + (DOT ? nlong : 0);
        double *partial = nullptr;
        if (DOT) {
            partial = (double *)scratch(sizeof(double) * ((size_t)kc * pstride + kc));
            if (!partial) return false;
        }
        {
            const double bytes = (sizeof(T) + 4.0) * A->nnz + 4.0 * (A->n + 1) + 2.0 * kc * (double)A->n * sizeof(T);
            const bool f64 = sizeof(T) == 8;
            ProfScope ps(DOT ? (f64 ? K_SPMM_DOT_F64 : K_SPMM_DOT_F32) : (f64 ? K_SPMM_F64 : K_SPMM_F32), bytes,
                         2.0 * kc * (double)A->nnz);
            if (f64)
                k_spmm<T, DOT, 4, false><<<nblk, kSpmmRows, 0, ctx().stream>>>(
                    A->n, A->rowptr, A->col, val, Xc, ldx, Yc, ldy, kc, g, partial, nblk, long_thr, pstride);
            else if (DOT)
                k_spmm<T, DOT, 4, true><<<((nblk + 7) / 8) * 8, kSpmmRows, 0, ctx().stream>>>(
                    A->n, A->rowptr, A->col, val, Xc, ldx, Yc, ldy, kc, g, partial, nblk, long_thr, pstride);
            else
                k_spmm<T, DOT, 8, true><<<((nblk + 7) / 8) * 8, kSpmmRows, 0, ctx().stream>>>(
                    A->n, A->rowptr, A->col, val, Xc, ldx, Yc, ldy, kc, g, partial, nblk, long_thr, pstride);
            if (nlong > 0)
                k_spmm_long<T, DOT><<<(nlong + kSpmmRows / 64 - 1) / (kSpmmRows / 64), kSpmmRows, 0,
                                      ctx().stream>>>(A->rowptr, A->col, val, Xc, ldx, Yc, ldy, kc, g, A->long_rows,
                                                      nlong, partial, pstride, nblk);
        }
        if (!check(hipGetLastError(), "k_spmm")) return false;
        if (DOT && !reduce_to_host(partial, kc, pstride, out + c0)) return false;
    }
    return true;
}

void launch_spmm(int prec, HIP_MAT *A, const void *X, int64_t ldx, void *Y, int64_t ldy, int k, const int *xidx)
{
    if (k <= 0 || failed()) return;
    if (prec == GCGE_HIP_FP64) spmm_t<double, false>(A, X, ldx, Y, ldy, k, xidx, nullptr);
    else spmm_t<float, false>(A, X, ldx, Y, ldy, k, xidx, nullptr);
}

bool spmm_dot(int prec, HIP_MAT *A, const void *X, int64_t ldx, void *Y, int64_t ldy, int k, double *out)
{
    if (k <= 0) return true;
    if (failed()) return false;
    if (prec == GCGE_HIP_FP64) return spmm_t<double, true>(A, X, ldx, Y, ldy, k, nullptr, out);
    return spmm_t<float, true>(A, X, ldx, Y, ldy, k, nullptr, out);
}

// Residual sums of kc <= kMaxCols columns on the row-major gather path;
// out[3j + q]: r.r, x.x, Bx.Bx of column j.  B, when given, has no row longer
// than kLongRow.
template <typename T, int G>
static bool residual_rows_g(HIP_MAT *A, HIP_MAT *B, const T *X, int64_t ldx, ColScalars g, int kc, double *out)
{
    const int n = A->n, kp = G < 64 ? G : ((kc + 63) / 64) * 64;
    T *Xt = (T *)row_stage(sizeof(T) * (size_t)n * kp);
    if (!Xt) return false;
    const int RWG = rows_per_wg(G);
    const int nblk = (n + RWG - 1) / RWG, nlong = A->nlong, nlb = (nlong + 3) / 4;
    const int pstride = nblk + nlong;
    double *partial = (double *)scratch(sizeof(double) * ((size_t)3 * kc * pstride + 3 * kc));
    if (!partial) return false;
    {
        ProfScope ps(K_RESIDUAL,
                     12.0 * A->nnz + 4.0 * (n + 1) + (B ? 12.0 * B->nnz + 4.0 * (n + 1) : 0.0) +
                         (double)kc * n * sizeof(T),
                     2.0 * kc * (double)(A->nnz + (B ? B->nnz : 0)));
        k_rows_from_cols<T><<<dim3((n + 63) / 64, (kp + 63) / 64), 256, 0, ctx().stream>>>(n, X, ldx, kc, g, kp,
                                                                                           Xt);
        const dim3 grid(nlb + nblk, G < 64 ? 1 : kp / 64);
        const int lthr = nlong > 0 ? kLongRow : INT_MAX;
        if (B)
            k_residual_rows<T, G, true><<<grid, 256, 0, ctx().stream>>>(
                n, A->rowptr, A->col, A->val64, B->rowptr, B->col, B->val64, Xt, kp, kc, g, A->long_rows, nlong,
                nlb, lthr, partial, pstride, nblk);
        else
            k_residual_rows<T, G, false><<<grid, 256, 0, ctx().stream>>>(
                n, A->rowptr, A->col, A->val64, nullptr, nullptr, nullptr, Xt, kp, kc, g, A->long_rows, nlong, nlb,
                lthr, partial, pstride, nblk);
    }
    if (!check(hipGetLastError(), "k_residual_rows")) return false;
    return reduce_to_host(partial, 3 * kc, pstride, out);
}

template <typename T>
static bool residual_rows(HIP_MAT *A, HIP_MAT *B, const T *X, int64_t ldx, ColScalars g, int kc, double *out)
{
    if (kc > 32) return residual_rows_g<T, 64>(A, B, X, ldx, g, kc, out);
    if (kc > 16) return residual_rows_g<T, 32>(A, B, X, ldx, g, kc, out);
    if (kc > 8) return residual_rows_g<T, 16>(A, B, X, ldx, g, kc, out);
    if (kc > 4) return residual_rows_g<T, 8>(A, B, X, ldx, g, kc, out);
    if (kc > 2) return residual_rows_g<T, 4>(A, B, X, ldx, g, kc, out);
    return residual_rows_g<T, 2>(A, B, X, ldx, g, kc, out);
}

template <typename T>
static bool residual_t(HIP_MAT *A, HIP_MAT *B, const void *X, int64_t ldx, const int *idx, int k,
                       const double *lambda, double *rnorm, double *xnorm, double *bxnorm)
{
    const int nblk = (A->n + kSpmmRows - 1) / kSpmmRows;
    std::vector<double> tmp(3 * kMaxCols);
    for (int c0 = 0; c0 < k; c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        ColScalars g = make_scalars(kc, lambda + c0, nullptr, idx + c0);
        // row gather when either matrix prefers it (B's rows must all be short)
        if ((A->gather_rows || (B && B->gather_rows)) && (!B || B->nlong == 0) && kc >= 2) {
            if (!residual_rows<T>(A, B, (const T *)X, ldx, g, kc, tmp.data())) return false;
            for (int j = 0; j < kc; j++) {
                rnorm[c0 + j] = std::sqrt(tmp[3 * j + 0]);
                xnorm[c0 + j] = std::sqrt(tmp[3 * j + 1]);
                if (bxnorm) bxnorm[c0 + j] = std::sqrt(tmp[3 * j + 2]);
            }
            continue;
        }
        double *partial = (double *)scratch(sizeof(double) * ((size_t)3 * kc * nblk + 3 * kc));
        if (!partial) return false;
        const double rbytes = 12.0 * A->nnz + 4.0 * (A->n + 1) + (B ? 12.0 * B->nnz + 4.0 * (B->n + 1) : 0.0) +
                              (double)kc * A->n * sizeof(T);
        {
            ProfScope ps(K_RESIDUAL, rbytes, 2.0 * kc * (double)(A->nnz + (B ? B->nnz : 0)));
            if (B)
                k_residual<T, true><<<nblk, kSpmmRows, 0, ctx().stream>>>(
                    A->n, A->rowptr, A->col, A->val64, B->rowptr, B->col, B->val64, (const T *)X, ldx, kc, g,
                    partial);
            else
                k_residual<T, false><<<nblk, kSpmmRows, 0, ctx().stream>>>(
                    A->n, A->rowptr, A->col, A->val64, nullptr, nullptr, nullptr, (const T *)X, ldx, kc, g,
                    partial);
        }
        if (!check(hipGetLastError(), "k_residual")) return false;
        if (!reduce_to_host(partial, 3 * kc, nblk, tmp.data())) return false;
        for (int j = 0; j < kc; j++) {
            rnorm[c0 + j] = std::sqrt(tmp[3 * j + 0]);
            xnorm[c0 + j] = std::sqrt(tmp[3 * j + 1]);
            if (bxnorm) bxnorm[c0 + j] = std::sqrt(tmp[3 * j + 2]);
        }
    }
    return true;
}

bool residual_norms(int prec, HIP_MAT *A, HIP_MAT *B, const void *X, int64_t ldx, const int *idx, int k,
                    const double *lambda, double *rnorm, double *xnorm, double *bxnorm)
{
    if (k <= 0) return true;
    if (failed()) return false;
    if (prec == GCGE_HIP_FP64) return residual_t<double>(A, B, X, ldx, idx, k, lambda, rnorm, xnorm, bxnorm);
    return residual_t<float>(A, B, X, ldx, idx, k, lambda, rnorm, xnorm, bxnorm);
}

bool cg_update(int prec, int64_t n, void *x, int64_t ldx, const void *p, int64_t ldp, const void *w,
               int64_t ldw, void *r, int64_t ldr, int k, const double *alpha, double *rr)
{
    if (k <= 0) return true;
    if (failed()) return false;
    const int nb = grid_1d(n, 256);
    for (int c0 = 0; c0 < k; c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        ColScalars s = make_scalars(kc, alpha + c0, nullptr, nullptr);
        double *partial = (double *)scratch(sizeof(double) * ((size_t)kc * nb + kc));
        if (!partial) return false;
        dim3 grid(nb, kc);
        {
            ProfScope ps(prec == GCGE_HIP_FP64 ? K_CG_UPDATE_F64 : K_CG_UPDATE_F32,
                         6.0 * kc * (double)n * elem_size(prec), 6.0 * kc * (double)n);
            if (prec == GCGE_HIP_FP64)
                k_cg_update<double><<<grid, kBlock, 0, ctx().stream>>>(
                    n, (double *)x + (int64_t)c0 * ldx, ldx, (const double *)p + (int64_t)c0 * ldp, ldp,
                    (const double *)w + (int64_t)c0 * ldw, ldw, (double *)r + (int64_t)c0 * ldr, ldr, s, partial);
            else
                k_cg_update<float><<<grid, kBlock, 0, ctx().stream>>>(
                    n, (float *)x + (int64_t)c0 * ldx, ldx, (const float *)p + (int64_t)c0 * ldp, ldp,
                    (const float *)w + (int64_t)c0 * ldw, ldw, (float *)r + (int64_t)c0 * ldr, ldr, s, partial);
        }
        if (!check(hipGetLastError(), "k_cg_update")) return false;
        if (!reduce_to_host(partial, kc, nb, rr + c0)) return false;
    }
    return true;
}

bool cg_resid(int prec, int64_t n, const void *w, int64_t ldw, void *r, int64_t ldr, int k, const double *alpha,
              double *rr)
{
    if (k <= 0) return true;
    if (failed()) return false;
    const int nb = grid_1d(n, 256);
    for (int c0 = 0; c0 < k; c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        ColScalars s = make_scalars(kc, alpha + c0, nullptr, nullptr);
        double *partial = (double *)scratch(sizeof(double) * ((size_t)kc * nb + kc));
        if (!partial) return false;
        dim3 grid(nb, kc);
        {
            ProfScope ps(prec == GCGE_HIP_FP64 ? K_CG_UPDATE_F64 : K_CG_UPDATE_F32,
                         3.0 * kc * (double)n * elem_size(prec), 4.0 * kc * (double)n);
            if (prec == GCGE_HIP_FP64)
                k_cg_resid<double><<<grid, kBlock, 0, ctx().stream>>>(
                    n, (const double *)w + (int64_t)c0 * ldw, ldw, (double *)r + (int64_t)c0 * ldr, ldr, s, partial);
            else
                k_cg_resid<float><<<grid, kBlock, 0, ctx().stream>>>(
                    n, (const float *)w + (int64_t)c0 * ldw, ldw, (float *)r + (int64_t)c0 * ldr, ldr, s, partial);
        }
        if (!check(hipGetLastError(), "k_cg_resid")) return false;
        if (!reduce_to_host(partial, kc, nb, rr + c0)) return false;
    }
    return true;
}

bool pcg_resid(int prec, int64_t n, const void *w, int64_t ldw, void *r, int64_t ldr, const void *dinv, int k,
               const double *alpha, double *rz, double *rr)
{
    if (k <= 0) return true;
    if (failed()) return false;
    const int nb = grid_1d(n, 256);
    std::vector<double> tmp(2 * kMaxCols);
    for (int c0 = 0; c0 < k; c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        ColScalars s = make_scalars(kc, alpha + c0, nullptr, nullptr);
        double *partial = (double *)scratch(sizeof(double) * ((size_t)2 * kc * nb + 2 * kc));
        if (!partial) return false;
        dim3 grid(nb, kc);
        {
            ProfScope ps(prec == GCGE_HIP_FP64 ? K_CG_UPDATE_F64 : K_CG_UPDATE_F32,
                         3.0 * kc * (double)n * elem_size(prec) + (double)n * elem_size(prec), 6.0 * kc * (double)n);
            if (prec == GCGE_HIP_FP64)
                k_pcg_resid<double><<<grid, kBlock, 0, ctx().stream>>>(
                    n, (const double *)w + (int64_t)c0 * ldw, ldw, (double *)r + (int64_t)c0 * ldr, ldr,
                    (const double *)dinv, s, partial);
            else
                k_pcg_resid<float><<<grid, kBlock, 0, ctx().stream>>>(
                    n, (const float *)w + (int64_t)c0 * ldw, ldw, (float *)r + (int64_t)c0 * ldr, ldr,
                    (const float *)dinv, s, partial);
        }
        if (!check(hipGetLastError(), "k_pcg_resid")) return false;
        if (!reduce_to_host(partial, 2 * kc, nb, tmp.data())) return false;
        for (int j = 0; j < kc; j++) {
            rz[c0 + j] = tmp[2 * j];
            rr[c0 + j] = tmp[2 * j + 1];
        }
    }
    return true;
}

void launch_diag_inv(HIP_MAT *A)
{
    k_diag_inv<<<(A->n + kBlock - 1) / kBlock, kBlock, 0, ctx().stream>>>(A->n, A->rowptr, A->col, A->val64,
                                                                         A->dinv64, A->dinv32);
    check(hipGetLastError(), "k_diag_inv");
}

void launch_cg_search(int prec, int64_t n, void *x, int64_t ldx, void *p, int64_t ldp, const void *r, int64_t ldr,
                      int kx, int kp, const double *alpha, const double *beta, const void *dinv)
{
    const int k = std::max(kx, kp);
    if (k <= 0 || failed()) return;
    const int nb = grid_1d(n, 1024);
    for (int c0 = 0; c0 < k; c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        const int kxc = std::max(0, std::min(kc, kx - c0)), kpc = std::max(0, std::min(kc, kp - c0));
        std::vector<double> aa(kc, 0.0), bb(kc, 0.0);
        for (int j = 0; j < kxc; j++) aa[j] = alpha[c0 + j];
        for (int j = 0; j < kpc; j++) bb[j] = beta[c0 + j];
        ColScalars s = make_scalars(kc, aa.data(), bb.data(), nullptr);
        dim3 grid(nb, kc);
        {
            ProfScope ps(prec == GCGE_HIP_FP64 ? K_CG_SEARCH_F64 : K_CG_SEARCH_F32,
                         (3.0 * kxc + 3.0 * kpc - 1.0 * std::min(kxc, kpc)) * (double)n * elem_size(prec) +
                             (dinv && kpc > 0 ? (double)n * elem_size(prec) : 0.0),
                         0.0);
#define GH_SEARCH(T_, PRE_)                                                                                     \
    k_cg_search<T_, PRE_><<<grid, kBlock, 0, ctx().stream>>>(n, (T_ *)x + (int64_t)c0 * ldx, ldx,              \
                                                              (T_ *)p + (int64_t)c0 * ldp, ldp,              \
                                                              (const T_ *)r + (int64_t)c0 * ldr, ldr, kxc, kpc, \
                                                              s, (const T_ *)dinv)
            if (prec == GCGE_HIP_FP64) {
                if (dinv) GH_SEARCH(double, true);
                else GH_SEARCH(double, false);
            } else {
                if (dinv) GH_SEARCH(float, true);
                else GH_SEARCH(float, false);
            }
#undef GH_SEARCH
        }
        if (!check(hipGetLastError(), "k_cg_search")) return;
    }
}

bool sampled_outer(int prec, HIP_MAT *A, const void *X, int64_t ldx, int k, const double *w, double *g_dev)
{
    if (failed()) return false;
    const int *rowidx = mat_rowidx(A);
    if (!rowidx) return false;
    if (!check(hipMemsetAsync(g_dev, 0, sizeof(double) * A->nnz, ctx().stream), "memset g")) return false;
    const int nb = (A->nnz + kBlock - 1) / kBlock;
    for (int c0 = 0; c0 < k; c0 += kMaxCols) {
        const int kc = std::min(kMaxCols, k - c0);
        ColScalars s = make_scalars(kc, w + c0, nullptr, nullptr);
        if (prec == GCGE_HIP_FP64)
            k_sampled_outer<double><<<nb, kBlock, 0, ctx().stream>>>(A->nnz, rowidx, A->col,
                                                                     (const double *)X + (int64_t)c0 * ldx, ldx,
                                                                     kc, s, g_dev);
        else
            k_sampled_outer<float><<<nb, kBlock, 0, ctx().stream>>>(A->nnz, rowidx, A->col,
                                                                    (const float *)X + (int64_t)c0 * ldx, ldx,
                                                                    kc, s, g_dev);
        if (!check(hipGetLastError(), "k_sampled_outer")) return false;
    }
    return true;
}

void launch_csr_gather_sum(int nparam, const int *jrowptr, const int *jnz, const double *jval, const double *g,
                           double *grad, double scale, bool accumulate)
{
    if (nparam <= 0 || failed()) return;
    k_csr_gather_sum<<<(nparam + kBlock - 1) / kBlock, kBlock, 0, ctx().stream>>>(nparam, jrowptr, jnz, jval, g,
                                                                                   grad, scale, accumulate ? 1 : 0);
    check(hipGetLastError(), "k_csr_gather_sum");
}

// matrix helpers used by hip_backend.cpp
void launch_rowidx(HIP_MAT *A)
{
    k_rowidx<<<(A->n + kBlock - 1) / kBlock, kBlock, 0, ctx().stream>>>(A->n, A->rowptr, A->rowidx);
    check(hipGetLastError(), "k_rowidx");
}

void launch_to_f32(int64_t n, const double *a, float *b)
{
    k_to_f32<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, ctx().stream>>>(n, a, b);
    check(hipGetLastError(), "k_to_f32");
}

}  // namespace gh
