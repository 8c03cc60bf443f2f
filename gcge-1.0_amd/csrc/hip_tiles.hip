// hip_tiles.hip -- Gram products G = V^T W and block linear combinations
// Y = beta Y + alpha X C on the gfx950 matrix cores.
//
// Both work on column-major multivectors (leading dimension a multiple of
// 256 elements, so every column is 2 KiB / 1 KiB aligned) and use the
// 16x16x4 MFMA shapes: v_mfma_f64_16x16x4f64 for FP64 data and
// v_mfma_f32_16x16x4f32 for FP32 data (FP32 accumulation).  Lane layout of
// these instructions, lane l = 0..63 of a wavefront:
//   A operand (16 x 4):  A[l % 16][l / 16]
//   B operand (4 x 16):  B[l / 16][l % 16]
//   D (16 x 16), slot t of lane l:  FP32  D[4 (l / 16) + t][l % 16]
//                                   FP64  D[4 t + l / 16][l % 16]
// (the FP64 result rows are interleaved; checked on the device).
//
// Gram: the n rows are cut into chunks; a workgroup owns one chunk and one
// output tile, and the reduction index k of the MFMA runs over rows.  Each
// lane loads R consecutive rows of one column as one aligned vector (32
// bytes), so a superstep of the wavefront covers 4 R rows and MFMA step t
// takes row  base + 4 R s + (l / 16) R + t  for both operands.  Partial
// tiles go to a per-chunk buffer that a second kernel sums in a fixed order.
// Very thin products (p <= 4, q <= 8 or the transpose) take a vector kernel
// that keeps the whole p x q block in registers and reads V and W once.
//
// LinComb: a wavefront owns 32 rows of Y (with more than 64 output columns,
// two wavefronts share the rows and split the columns) and walks
// its output columns in blocks of 16; X supplies the A operand straight from HBM (16 consecutive
// rows per lane group), C the B operand from a copy packed on the host in
// operand order.  Thin combinations (p, q <= 8) take a vector kernel with C
// in the kernel arguments and 16-byte row vectors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hip_common.h"

namespace gh {

namespace {

constexpr int kThreads = 256;       // 4 wavefronts per workgroup
constexpr int kGramTargetWG = 2048; // workgroups a Gram launch aims for
constexpr int kGramSlabWG = 512;    // workgroups of a slab Gram launch (two per CU)

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));
typedef float fx8 __attribute__((ext_vector_type(8)));

template <typename T>
struct MT;

template <>
struct MT<double> {
    using acc_t = dx4;
    using vec_t = dx4;              // R rows of one column
    static constexpr int R = 4;
    __device__ static __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // result row of slot t in lane group g = l / 16
    __device__ static __forceinline__ int drow(int g, int t) { return 4 * t + g; }
};

template <>
struct MT<float> {
    using acc_t = fx4;
    using vec_t = fx8;
    static constexpr int R = 8;
    __device__ static __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static __forceinline__ int drow(int g, int t) { return 4 * g + t; }
};

// R consecutive rows [row, row + R) of one column, zero past `end`
template <typename T>
__device__ __forceinline__ void load_rows(const T *__restrict__ col, int64_t row, int64_t end,
                                          T (&v)[MT<T>::R])
{
    constexpr int R = MT<T>::R;
    if (row + R <= end) {
        const typename MT<T>::vec_t x = *reinterpret_cast<const typename MT<T>::vec_t *>(col + row);
#pragma unroll
        for (int t = 0; t < R; t++) v[t] = x[t];
    } else {
#pragma unroll
        for (int t = 0; t < R; t++) v[t] = (row + t < end) ? col[row + t] : T(0);
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Workgroups are dealt round-robin to the 8 XCDs; this renumbering gives
// each XCD a contiguous range of logical blocks, so the tiles of one row
// chunk run together on one XCD and share its L2.
__device__ __forceinline__ int xcd_logical_block()
{
    const int per = (int)(gridDim.x >> 3);
    return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// ---------------------------------------------------------------- Gram (MFMA)
// Workgroup tile TP x TQ = (GP WP 16) x (GQ WQ 16); its 4 wavefronts are
// arranged GP x GQ over the tile and GK = 4 / (GP GQ) ways over the rows.
// partial[chunk][j * p + i] = sum over the chunk's rows of V(r, i) W(r, j).
template <typename T, int WP, int WQ, int GP, int GQ>
__global__ __launch_bounds__(kThreads) void k_gram_mfma(int64_t n, const T *__restrict__ V, int64_t ldv, int p,
                                                        const T *__restrict__ W, int64_t ldw, int q, int tiles_p,
                                                        int tiles_q, int nchunk, int64_t rows_per_chunk,
                                                        int upper_only, double *__restrict__ partial)
{
    constexpr int GK = 4 / (GP * GQ);
    constexpr int R = MT<T>::R;
    constexpr int S = 4 * R;  // rows per wavefront superstep
    constexpr int TP = GP * WP * 16, TQ = GQ * WQ * 16;
    using acc_t = typename MT<T>::acc_t;
    __shared__ double red[GK > 1 ? (GK - 1) * GP * GQ * WP * WQ * 4 * 64 : 1];

    const int ntile = tiles_p * tiles_q;
    const int logical = xcd_logical_block();
    if (logical >= ntile * nchunk) return;
    const int chunk = logical / ntile, tile = logical % ntile;
    const int i0t = (tile % tiles_p) * TP, j0t = (tile / tiles_p) * TQ;
    if (upper_only && i0t > j0t + TQ - 1) return;  // tile entirely below the diagonal

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wp = wave % GP, wq = (wave / GP) % GQ, wk = wave / (GP * GQ);
    const int i0 = i0t + wp * WP * 16, j0 = j0t + wq * WQ * 16;
    const int m = lane & 15, kk = lane >> 4;
    const int64_t rb = (int64_t)chunk * rows_per_chunk;
    const int64_t re = min(n, rb + rows_per_chunk);

    // Columns past p (q) are clamped to the last one: they feed only result
    // rows (columns) >= p (q), which are never stored, so no lane branches.
    const T *vc[WP];
    const T *wc[WQ];
#pragma unroll
    for (int x = 0; x < WP; x++) vc[x] = V + (int64_t)min(i0 + x * 16 + m, p - 1) * ldv;
#pragma unroll
    for (int y = 0; y < WQ; y++) wc[y] = W + (int64_t)min(j0 + y * 16 + m, q - 1) * ldw;

    acc_t acc[WP][WQ];
#pragma unroll
    for (int x = 0; x < WP; x++)
#pragma unroll
        for (int y = 0; y < WQ; y++) acc[x][y] = acc_t{};

    // full supersteps (every row below re): unguarded vector loads, the next
    // superstep's operands loaded before this one's MFMAs are issued
    using vec_t = typename MT<T>::vec_t;
    const int64_t step = (int64_t)GK * S;
    const int64_t full_end = rb + ((re - rb) / S) * S;
    int64_t s = rb + (int64_t)wk * S;
    if (s < full_end) {
        auto load = [&](int64_t ss, vec_t (&aa)[WP], vec_t (&bb)[WQ]) {
#pragma unroll
            for (int x = 0; x < WP; x++) aa[x] = *reinterpret_cast<const vec_t *>(vc[x] + ss + (int64_t)kk * R);
#pragma unroll
            for (int y = 0; y < WQ; y++) bb[y] = *reinterpret_cast<const vec_t *>(wc[y] + ss + (int64_t)kk * R);
        };
        auto mma = [&](const vec_t (&aa)[WP], const vec_t (&bb)[WQ]) {
#pragma unroll
            for (int t = 0; t < R; t++)
#pragma unroll
                for (int x = 0; x < WP; x++)
#pragma unroll
                    for (int y = 0; y < WQ; y++) acc[x][y] = MT<T>::mma(aa[x][t], bb[y][t], acc[x][y]);
        };
        if constexpr (GP * GQ == 4) {
            // 64 x 64 tiles: one operand buffer (register budget for four
            // wavefronts per SIMD); latency is hidden by occupancy
            for (; s < full_end; s += step) {
                vec_t a0[WP], b0[WQ];
                load(s, a0, b0);
                mma(a0, b0);
            }
        } else {
            // supersteps in pairs with two operand buffers (loads of one in
            // flight while the MFMAs of the other are issued)
            vec_t a0[WP], b0[WQ], a1[WP], b1[WQ];
            // the last full superstep of this wavefront
            const int64_t last_full = s + ((full_end - 1 - s) / step) * step;
            load(s, a0, b0);
            for (; s + step < full_end; s += 2 * step) {
                load(s + step, a1, b1);
                mma(a0, b0);
                // unconditional (clamped to a full superstep) so the wait
                // before the next MFMAs covers only the older loads
                load(min(s + 2 * step, last_full), a0, b0);
                mma(a1, b1);
            }
            if (s < full_end) {
                mma(a0, b0);
                s += step;
            }
        }
    }
    // the partial superstep at the end of the chunk (rows past re read as 0)
    if (s < re) {
        const int64_t row = s + (int64_t)kk * R;
        T a[WP][R], b[WQ][R];
#pragma unroll
        for (int x = 0; x < WP; x++) load_rows<T>(vc[x], row, re, a[x]);
#pragma unroll
        for (int y = 0; y < WQ; y++) load_rows<T>(wc[y], row, re, b[y]);
#pragma unroll
        for (int t = 0; t < R; t++)
#pragma unroll
            for (int x = 0; x < WP; x++)
#pragma unroll
                for (int y = 0; y < WQ; y++) acc[x][y] = MT<T>::mma(a[x][t], b[y][t], acc[x][y]);
    }

    // sum the GK row-split partial tiles in wavefront order: the wavefronts
    // of row slice wk > 0 park their tiles in red[((wk - 1) GP GQ + wpq) SLOTS
    // + slot][lane], wpq = wp + GP wq, and slice 0 adds them
    const int wpq = wp + GP * wq;
    if constexpr (GK > 1) {
        constexpr int SLOTS = WP * WQ * 4;
        if (wk > 0) {
#pragma unroll
            for (int x = 0; x < WP; x++)
#pragma unroll
                for (int y = 0; y < WQ; y++)
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        red[(((size_t)(wk - 1) * GP * GQ + wpq) * SLOTS + (x * WQ + y) * 4 + t) * 64 + lane] =
                            (double)acc[x][y][t];
        }
        __syncthreads();
        if (wk > 0) return;
    }

    double *out = partial + (int64_t)chunk * p * q;
#pragma unroll
    for (int x = 0; x < WP; x++)
#pragma unroll
        for (int y = 0; y < WQ; y++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                double v = (double)acc[x][y][t];
                if constexpr (GK > 1) {
                    constexpr int SLOTS = WP * WQ * 4;
#pragma unroll
                    for (int w = 0; w < GK - 1; w++)
                        v += red[(((size_t)w * GP * GQ + wpq) * SLOTS + (x * WQ + y) * 4 + t) * 64 + lane];
                }
                const int i = i0 + x * 16 + MT<T>::drow(kk, t), j = j0 + y * 16 + m;
                if (i < p && j < q) out[(int64_t)j * p + i] = v;
            }
}

// ---------------------------------------------------------------- Gram (LDS slabs)
// Workgroup output block (MB 16) x (NB 16), NB <= 8 at run time, over one
// row chunk.  Every SLAB rows (128 bytes of each column) the 256 threads copy
// the block's columns of V and W into LDS (column-major, 16-byte pad per
// column: 144-byte column stride, so neither the 16-byte copies nor the
// operand reads of 16 consecutive columns meet a bank twice), two buffers
// deep; wavefront w runs the MFMAs of tile columns w and w + 4 for all MB
// tile rows.  V and W thus leave memory once per block instead of once per
// 64 x 64 tile, and p, q are padded to 16 rather than 64.
template <typename T, int MB>
__global__ __launch_bounds__(kThreads) void k_gram_slab(int64_t n, const T *__restrict__ V, int64_t ldv, int p,
                                                        const T *__restrict__ W, int64_t ldw, int q, int NB,
                                                        int blocks_p, int blocks_q, int nchunk,
                                                        int64_t rows_per_chunk, int upper_only,
                                                        double *__restrict__ partial)
{
    constexpr int SLAB = 128 / (int)sizeof(T);     // rows per stage
    constexpr int LD = SLAB + 16 / (int)sizeof(T);  // LDS column stride (elements)
    constexpr int TP = MB * 16;
    constexpr int COLS = TP + 128;                  // V block + up to 8 W tiles
    constexpr int UPT = (COLS * 8 + kThreads - 1) / kThreads;  // 16-byte units per thread
    constexpr int E = 16 / (int)sizeof(T);          // elements per unit
    using acc_t = typename MT<T>::acc_t;
    typedef T u16 __attribute__((ext_vector_type(E)));
    __shared__ __attribute__((aligned(16))) T sbuf[2][COLS * LD];

    const int nblock = blocks_p * blocks_q;
    const int logical = xcd_logical_block();
    if (logical >= nblock * nchunk) return;
    const int chunk = logical / nblock, blk = logical % nblock;
    const int i0 = (blk % blocks_p) * TP, j0 = (blk / blocks_p) * NB * 16;
    const int64_t rb = (int64_t)chunk * rows_per_chunk;
    const int64_t re = min(n, rb + rows_per_chunk);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kk = lane >> 4;
    const int ncols = TP + NB * 16;

    // this thread's copy units: column c (V for c < TP, else W), 16 bytes
    // at row offset E * seg of the slab; columns past p (q) repeat the last
    // one and feed only result entries that are never stored
    const T *src[UPT];
    int dst[UPT];
#pragma unroll
    for (int u = 0; u < UPT; u++) {
        const int unit = threadIdx.x + u * kThreads;
        const int c = unit >> 3, seg = unit & 7;
        if (c < ncols) {
            src[u] = (c < TP ? V + (int64_t)min(i0 + c, p - 1) * ldv
                             : W + (int64_t)min(j0 + c - TP, q - 1) * ldw) + seg * E;
            dst[u] = c * LD + seg * E;
        } else {
            src[u] = nullptr;
            dst[u] = -1;
        }
    }
    u16 reg[UPT];
    auto fetch = [&](int64_t r) {
        if (r + SLAB <= re) {
#pragma unroll
            for (int u = 0; u < UPT; u++)
                if (dst[u] >= 0) reg[u] = *reinterpret_cast<const u16 *>(src[u] + r);
        } else {
#pragma unroll
            for (int u = 0; u < UPT; u++)
                if (dst[u] >= 0) {
                    const int64_t r0 = r + dst[u] % LD;  // row of the unit's first element
#pragma unroll
                    for (int e = 0; e < E; e++) reg[u][e] = (r0 + e < re) ? src[u][r + e] : T(0);
                }
        }
    };
    auto park = [&](int b) {
#pragma unroll
        for (int u = 0; u < UPT; u++)
            if (dst[u] >= 0) *reinterpret_cast<u16 *>(&sbuf[b][dst[u]]) = reg[u];
    };

    acc_t acc[MB][2];
#pragma unroll
    for (int x = 0; x < MB; x++) acc[x][0] = acc[x][1] = acc_t{};
    // tiles this wavefront computes: column y = wave + 4 yy; with upper_only,
    // tiles entirely below the diagonal are skipped
    bool live[MB][2];
#pragma unroll
    for (int yy = 0; yy < 2; yy++) {
        const int y = wave + 4 * yy;
#pragma unroll
        for (int x = 0; x < MB; x++)
            live[x][yy] = y < NB && !(upper_only && i0 + x * 16 > j0 + y * 16 + 15);
    }

    const int nslab = (int)((re - rb + SLAB - 1) / SLAB);
    if (nslab > 0) {
        fetch(rb);
        park(0);
    }
    __syncthreads();
    for (int sl = 0; sl < nslab; sl++) {
        const int b = sl & 1;
        const bool more = sl + 1 < nslab;
        if (more) fetch(rb + (int64_t)(sl + 1) * SLAB);
        const T *sv = sbuf[b];
        const T *sw = sbuf[b] + TP * LD;
#pragma unroll
        for (int ks = 0; ks < SLAB / 4; ks++) {
            const int kr = ks * 4 + kk;
            T a[MB], bq[2];
#pragma unroll
            for (int x = 0; x < MB; x++) a[x] = sv[(x * 16 + m) * LD + kr];
#pragma unroll
            for (int yy = 0; yy < 2; yy++) bq[yy] = sw[((wave + 4 * yy) * 16 + m) * LD + kr];
#pragma unroll
            for (int yy = 0; yy < 2; yy++)
#pragma unroll
                for (int x = 0; x < MB; x++)
                    if (live[x][yy]) acc[x][yy] = MT<T>::mma(a[x], bq[yy], acc[x][yy]);
        }
        if (more) park(b ^ 1);
        __syncthreads();
    }

    double *out = partial + (int64_t)chunk * p * q;
#pragma unroll
    for (int yy = 0; yy < 2; yy++) {
        const int y = wave + 4 * yy;
        if (y >= NB) continue;
#pragma unroll
        for (int x = 0; x < MB; x++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int i = i0 + x * 16 + MT<T>::drow(kk, t), j = j0 + y * 16 + m;
                if (i < p && j < q) out[(int64_t)j * p + i] = (double)acc[x][yy][t];
            }
    }
}

// ---------------------------------------------------------------- Gram (thin)
// p <= PM, q <= QM; every thread keeps the PM x QM block of its rows in
// registers.  partial[block][o] with o = j * ldo + i, or i * ldo + j when
// transpose is set (the caller swapped V and W).
template <typename T, int PM, int QM>
__global__ __launch_bounds__(kThreads) void k_gram_thin(int64_t n, const T *__restrict__ V, int64_t ldv, int p,
                                                        const T *__restrict__ W, int64_t ldw, int q, int transpose,
                                                        int ldo, double *__restrict__ partial)
{
    __shared__ double sh[4][PM * QM];
    double acc[PM][QM];
#pragma unroll
    for (int a = 0; a < PM; a++)
#pragma unroll
        for (int b = 0; b < QM; b++) acc[a][b] = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kThreads;
#pragma unroll 2
    for (int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x; r < n; r += stride) {
        double v[PM], w[QM];
#pragma unroll
        for (int a = 0; a < PM; a++) v[a] = (a < p) ? (double)V[(int64_t)a * ldv + r] : 0.0;
#pragma unroll
        for (int b = 0; b < QM; b++) w[b] = (b < q) ? (double)W[(int64_t)b * ldw + r] : 0.0;
#pragma unroll
        for (int a = 0; a < PM; a++)
#pragma unroll
            for (int b = 0; b < QM; b++) acc[a][b] = fma(v[a], w[b], acc[a][b]);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int a = 0; a < PM; a++)
#pragma unroll
        for (int b = 0; b < QM; b++) {
            const double s = wave_sum(acc[a][b]);
            if (lane == 0) sh[wave][a * QM + b] = s;
        }
    __syncthreads();
    const int pq = p * q;
    for (int e = threadIdx.x; e < PM * QM; e += kThreads) {
        const int a = e / QM, b = e % QM;
        if (a < p && b < q) {
            const double s = (sh[0][e] + sh[1][e]) + (sh[2][e] + sh[3][e]);
            const int o = transpose ? a * ldo + b : b * ldo + a;
            partial[(int64_t)blockIdx.x * pq + o] = s;
        }
    }
}

// G[e] = sum_c partial[c][e]: 64 entries per workgroup, the 4 wavefronts take
// every 4th chunk, combined in a fixed order.  With upper_only, entries
// below the diagonal (i > j, e = j * p + i) are skipped.
__global__ __launch_bounds__(kThreads) void k_sum_chunks(const double *__restrict__ partial, int nchunk, int p,
                                                         int64_t pq, int upper_only, double *__restrict__ G)
{
    __shared__ double sh[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t e = (int64_t)blockIdx.x * 64 + lane;
    bool live = e < pq;
    if (live && upper_only) live = (e % p) <= (e / p);
    double acc = 0.0;
    if (live)
        for (int c = wave; c < nchunk; c += 4) acc += partial[(int64_t)c * pq + e];
    sh[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && e < pq) G[e] = live ? (sh[0][lane] + sh[1][lane]) + (sh[2][lane] + sh[3][lane]) : 0.0;
}

// ---------------------------------------------------------------- LinComb (MFMA)
// Cp[((kb * nqb) + jb) * 64 + l] = C(4 kb + l / 16, 16 jb + l % 16), zero
// past p and q.  A wavefront owns rows [rb, rb + 32) as two 16-row blocks and
// walks the output in groups of RN 16-column blocks; slot t of lane l holds
// row drow(l / 16, t) of its block and column l % 16.
// Result rows of one lane after regrouping: lane (g = l / 16, m) returns
// rows 4 g .. 4 g + 3 of column m of the 16 x 16 block.  For FP32 these are
// already its four slots; for FP64 (slot t holds row 4 t + g) the four lanes
// of a column exchange values in four rounds: in round k, lane g reads the
// slot g of lane (g + k) mod 4.
template <typename T>
__device__ __forceinline__ typename MT<T>::acc_t rows4(typename MT<T>::acc_t v, int g);

template <>
__device__ __forceinline__ fx4 rows4<float>(fx4 v, int)
{
    return v;
}

template <>
__device__ __forceinline__ dx4 rows4<double>(dx4 v, int g)
{
    const int lane = threadIdx.x & 63, m = lane & 15;
    dx4 w;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int src = (g + k) & 3;             // lane group this lane reads
        const int give = (g - k) & 3;            // slot this lane sends: the reader's g
        double send = v[0];
        send = give == 1 ? v[1] : send;
        send = give == 2 ? v[2] : send;
        send = give == 3 ? v[3] : send;
        const double got = __shfl(send, m + 16 * src, 64);
        w[0] = src == 0 ? got : w[0];
        w[1] = src == 1 ? got : w[1];
        w[2] = src == 2 ? got : w[2];
        w[3] = src == 3 ? got : w[3];
    }
    return w;
}

// y[0..3] = beta y + alpha v with the branch rules of y = a x + b y, as one
// 16- (FP32) or 32-byte (FP64) access; y is 4-element aligned
template <typename T>
__device__ __forceinline__ void store_rows4(T *y, typename MT<T>::acc_t v, T al, T be)
{
    typedef T v4 __attribute__((ext_vector_type(4)));
    v4 out;
    if (be == T(0)) {
#pragma unroll
        for (int t = 0; t < 4; t++) out[t] = (al == T(1)) ? (T)v[t] : al * (T)v[t];
    } else {
        const v4 old = *reinterpret_cast<const v4 *>(y);
#pragma unroll
        for (int t = 0; t < 4; t++) out[t] = be * old[t] + al * (T)v[t];
    }
    *reinterpret_cast<v4 *>(y) = out;
}

template <typename T, int RN, int WC>
__global__ __launch_bounds__(kThreads) void k_lincomb_mfma(int64_t n, const T *__restrict__ X, int64_t ldx, int p,
                                                           T *__restrict__ Y, int64_t ldy, int q,
                                                           const T *__restrict__ Cp, int nkb, int nqb, double alpha,
                                                           double beta)
{
    using acc_t = typename MT<T>::acc_t;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kk = lane >> 4;
    // WC wavefronts share one 32-row block and split its column groups
    const int64_t rb = ((int64_t)blockIdx.x * (4 / WC) + wave / WC) * 32;
    const int wcol = wave % WC;
    if (rb >= n) return;  // no workgroup barrier in this kernel
    // Rows past n are clamped to the last row: they feed only result rows
    // >= n, which are never stored.  nqb is a multiple of RN (Cp is padded
    // with zero blocks), and only the last k block can be partial.
    const int64_t ra0 = min(rb + m, n - 1), ra1 = min(rb + 16 + m, n - 1);
    const int nkb_full = p / 4;
    const T al = (T)alpha, be = (T)beta;
    for (int jb0 = wcol * RN; jb0 < nqb; jb0 += WC * RN) {
        acc_t acc0[RN], acc1[RN];
#pragma unroll
        for (int r = 0; r < RN; r++) {
            acc0[r] = acc_t{};
            acc1[r] = acc_t{};
        }
        if (nkb_full > 0) {
            // k blocks in pairs with two operand buffers: the loads of one
            // block are in flight while the MFMAs of the other are issued
            const T *xk = X + (int64_t)kk * ldx;
            const T *cb = Cp + (int64_t)jb0 * 64 + lane;
            const int64_t xs = (int64_t)4 * ldx, cs = (int64_t)nqb * 64;
            T a0[2], a1[2], b0[RN], b1[RN];
            auto load = [&](int kb, T (&aa)[2], T (&bb)[RN]) {
                const T *x = xk + (int64_t)kb * xs;
                const T *c = cb + (int64_t)kb * cs;
                aa[0] = x[ra0];
                aa[1] = x[ra1];
#pragma unroll
                for (int r = 0; r < RN; r++) bb[r] = c[r * 64];
            };
            auto mma = [&](const T (&aa)[2], const T (&bb)[RN]) {
#pragma unroll
                for (int r = 0; r < RN; r++) {
                    acc0[r] = MT<T>::mma(aa[0], bb[r], acc0[r]);
                    acc1[r] = MT<T>::mma(aa[1], bb[r], acc1[r]);
                }
            };
            load(0, a0, b0);
            int kb = 0;
            for (; kb + 1 < nkb_full; kb += 2) {
                load(kb + 1, a1, b1);
                mma(a0, b0);
                // unconditional (clamped) so the wait before the next MFMAs
                // covers only the older loads
                load(min(kb + 2, nkb_full - 1), a0, b0);
                mma(a1, b1);
            }
            if (kb < nkb_full) mma(a0, b0);
        }
        if (nkb_full < nkb) {
            // partial last k block: columns past p read as 0
            const int k = nkb_full * 4 + kk;
            const bool kok = k < p;
            const T *xk = X + (int64_t)(kok ? k : 0) * ldx;
            const T a0 = kok ? xk[ra0] : T(0);
            const T a1 = kok ? xk[ra1] : T(0);
            const T *cb = Cp + ((int64_t)nkb_full * nqb + jb0) * 64 + lane;
#pragma unroll
            for (int r = 0; r < RN; r++) {
                const T bb = cb[r * 64];
                acc0[r] = MT<T>::mma(a0, bb, acc0[r]);
                acc1[r] = MT<T>::mma(a1, bb, acc1[r]);
            }
        }
        // epilogue: lane (kk, m) ends with rows rb + 16 h + 4 kk .. + 3 of
        // column m in registers (the FP64 result rows are regrouped across
        // the four lanes of a column first), written as one vector
#pragma unroll
        for (int r = 0; r < RN; r++) {
            const int j = (jb0 + r) * 16 + m;
            T *yj = Y + (int64_t)min(j, q - 1) * ldy;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const acc_t v = rows4<T>(h == 0 ? acc0[r] : acc1[r], kk);
                const int64_t row = rb + 16 * h + 4 * kk;
                if (j >= q) continue;
                if (row + 4 <= n) {
                    store_rows4<T>(yj + row, v, al, be);
                } else {
#pragma unroll
                    for (int t = 0; t < 4; t++)
                        if (row + t < n) {
                            const T vt = (T)v[t];
                            yj[row + t] = (be == T(0)) ? ((al == T(1)) ? vt : al * vt) : be * yj[row + t] + al * vt;
                        }
                }
            }
        }
    }
}

// ---------------------------------------------------------------- LinComb (thin)
// p <= PM, q <= QM: C travels in the kernel arguments; each thread owns VR
// consecutive rows (one 16-byte vector per column) and forms
// y_j = beta y_j + alpha sum_i x_i C(i, j), the sum as fma from zero in i order.
struct ThinCoef {
    double c[64];  // C(i, j) at j * p + i
};

template <typename T, int PM, int QM>
__global__ __launch_bounds__(kThreads) void k_lincomb_thin(int64_t n, const T *__restrict__ X, int64_t ldx, int p,
                                                           T *__restrict__ Y, int64_t ldy, int q, ThinCoef cf,
                                                           double alpha, double beta)
{
    constexpr int VR = 16 / sizeof(T);
    typedef T vec_t __attribute__((ext_vector_type(VR)));
    T c[PM][QM];
#pragma unroll
    for (int i = 0; i < PM; i++)
#pragma unroll
        for (int j = 0; j < QM; j++) c[i][j] = (i < p && j < q) ? (T)cf.c[j * p + i] : T(0);
    const T al = (T)alpha, be = (T)beta;
    const int64_t stride = (int64_t)gridDim.x * kThreads * VR;
    for (int64_t r = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * VR; r < n; r += stride) {
        T x[PM][VR];
        const bool full = r + VR <= n;
#pragma unroll
        for (int i = 0; i < PM; i++) {
            if (i < p) {
                const T *xi = X + (int64_t)i * ldx + r;
                if (full) {
                    const vec_t v = *reinterpret_cast<const vec_t *>(xi);
#pragma unroll
                    for (int t = 0; t < VR; t++) x[i][t] = v[t];
                } else {
#pragma unroll
                    for (int t = 0; t < VR; t++) x[i][t] = (r + t < n) ? xi[t] : T(0);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < QM; j++) {
            if (j >= q) break;
            T *yj = Y + (int64_t)j * ldy + r;
            T out[VR];
#pragma unroll
            for (int t = 0; t < VR; t++) {
                T acc = T(0);
#pragma unroll
                for (int i = 0; i < PM; i++)
                    if (i < p) acc = fma(x[i][t], c[i][j], acc);
                out[t] = acc;
            }
            if (full) {
                vec_t y;
                if (be != T(0)) y = *reinterpret_cast<const vec_t *>(yj);
#pragma unroll
                for (int t = 0; t < VR; t++)
                    y[t] = (be == T(0)) ? ((al == T(1)) ? out[t] : al * out[t]) : be * y[t] + al * out[t];
                *reinterpret_cast<vec_t *>(yj) = y;
            } else {
#pragma unroll
                for (int t = 0; t < VR; t++)
                    if (r + t < n)
                        yj[t] = (be == T(0)) ? ((al == T(1)) ? out[t] : al * out[t]) : be * yj[t] + al * out[t];
            }
        }
    }
}

template <typename T>
void launch_lincomb_thin(int64_t n, const void *X, int64_t ldx, int p, void *Y, int64_t ldy, int q,
                         const ThinCoef &cf, double alpha, double beta)
{
    constexpr int VR = 16 / sizeof(T);
    const int64_t threads = (n + VR - 1) / VR;
    const int grid = (int)std::min<int64_t>(8192, std::max<int64_t>(1, (threads + kThreads - 1) / kThreads));
    const int pm = p <= 1 ? 1 : p <= 2 ? 2 : p <= 4 ? 4 : 8;
    const int qm = q <= 1 ? 1 : q <= 2 ? 2 : q <= 4 ? 4 : 8;
#define GH_LT(P_, Q_)                                                                                          \
    if (pm == P_ && qm == Q_) {                                                                                 \
        k_lincomb_thin<T, P_, Q_><<<grid, kThreads, 0, ctx().stream>>>(n, (const T *)X, ldx, p, (T *)Y, ldy, q, \
                                                                       cf, alpha, beta);                        \
        return;                                                                                                 \
    }
    GH_LT(1, 1) GH_LT(1, 2) GH_LT(1, 4) GH_LT(1, 8)
    GH_LT(2, 1) GH_LT(2, 2) GH_LT(2, 4) GH_LT(2, 8)
    GH_LT(4, 1) GH_LT(4, 2) GH_LT(4, 4) GH_LT(4, 8)
    GH_LT(8, 1) GH_LT(8, 2) GH_LT(8, 4) GH_LT(8, 8)
#undef GH_LT
}

template <typename T, int PM, int QM>
void launch_thin(int grid, int64_t n, const void *V, int64_t ldv, int p, const void *W, int64_t ldw, int q,
                 int transpose, int ldo, double *partial)
{
    k_gram_thin<T, PM, QM><<<grid, kThreads, 0, ctx().stream>>>(n, (const T *)V, ldv, p, (const T *)W, ldw, q,
                                                                transpose, ldo, partial);
}

template <typename T>
void launch_thin_sized(int grid, int64_t n, const void *V, int64_t ldv, int p, const void *W, int64_t ldw, int q,
                       int transpose, int ldo, double *partial)
{
    const int pm = p <= 1 ? 1 : p <= 2 ? 2 : 4;
    const int qm = q <= 1 ? 1 : q <= 2 ? 2 : q <= 4 ? 4 : 8;
#define GH_THIN(P_, Q_) \
    if (pm == P_ && qm == Q_) return launch_thin<T, P_, Q_>(grid, n, V, ldv, p, W, ldw, q, transpose, ldo, partial);
    GH_THIN(1, 1) GH_THIN(1, 2) GH_THIN(1, 4) GH_THIN(1, 8)
    GH_THIN(2, 1) GH_THIN(2, 2) GH_THIN(2, 4) GH_THIN(2, 8)
    GH_THIN(4, 1) GH_THIN(4, 2) GH_THIN(4, 4) GH_THIN(4, 8)
#undef GH_THIN
}

struct GramShape {
    int WP, WQ, GP, GQ;
};

GramShape gram_shape(int p, int q)
{
    if (p <= 16 && q <= 16) return {1, 1, 1, 1};  // 16 x 16 tiles, rows split 4 ways
    if (p <= 32 && q <= 32) return {2, 2, 1, 1};  // 32 x 32, rows split 4 ways
    if (q <= 16) return {2, 1, 2, 1};             // 64 x 16, rows split 2 ways
    if (p <= 16) return {1, 2, 1, 2};             // 16 x 64, rows split 2 ways
    return {2, 2, 2, 2};                          // 64 x 64
}

template <typename T>
void launch_gram_mfma(const GramShape &g, int grid, int64_t n, const void *V, int64_t ldv, int p, const void *W,
                      int64_t ldw, int q, int tp, int tq, int nchunk, int64_t rpc, int upper, double *partial)
{
#define GH_GRAM(WP_, WQ_, GP_, GQ_)                                                                             \
    if (g.WP == WP_ && g.WQ == WQ_ && g.GP == GP_ && g.GQ == GQ_) {                                             \
        k_gram_mfma<T, WP_, WQ_, GP_, GQ_><<<grid, kThreads, 0, ctx().stream>>>(                                \
            n, (const T *)V, ldv, p, (const T *)W, ldw, q, tp, tq, nchunk, rpc, upper, partial);                \
        return;                                                                                                 \
    }
    GH_GRAM(1, 1, 1, 1)
    GH_GRAM(2, 2, 1, 1)
    GH_GRAM(2, 1, 2, 1)
    GH_GRAM(1, 2, 1, 2)
    GH_GRAM(2, 2, 2, 2)
#undef GH_GRAM
}

template <typename T>
void launch_gram_slab(int MB, int grid, int64_t n, const void *V, int64_t ldv, int p, const void *W, int64_t ldw,
                      int q, int NB, int bp, int bq, int nchunk, int64_t rpc, int upper, double *partial)
{
#define GH_SLAB(MB_)                                                                                            \
    if (MB == MB_) {                                                                                            \
        k_gram_slab<T, MB_><<<grid, kThreads, 0, ctx().stream>>>(n, (const T *)V, ldv, p, (const T *)W, ldw, q, \
                                                                 NB, bp, bq, nchunk, rpc, upper, partial);      \
        return;                                                                                                 \
    }
    GH_SLAB(1) GH_SLAB(2) GH_SLAB(3) GH_SLAB(4) GH_SLAB(5) GH_SLAB(6) GH_SLAB(7) GH_SLAB(8)
#undef GH_SLAB
}

bool aligned32(const void *p) { return ((uintptr_t)p & 31u) == 0; }

}  // namespace

bool gram(int prec, int64_t n, const void *V, int64_t ldv, int p, const void *W, int64_t ldw, int q, double *G,
          int ldg, bool upper_only)
{
    if (p <= 0 || q <= 0) return true;
    if (failed()) return false;
    const size_t es = elem_size(prec);
    const int64_t pq = (int64_t)p * q;
    const bool thin = (p <= 4 && q <= 8) || (q <= 4 && p <= 8);
    // launch geometry and scratch first: growing the scratch synchronizes the
    // stream, which must not happen inside a profiled launch
    int nchunk = 0;
    int64_t rpc = 0;
    int tp = 1, tq = 1;
    GramShape g{1, 1, 1, 1};
    // slab kernel beyond 64 columns on either side (blocks of up to 8 x 8
    // tiles); up to 64 x 64 the single-tile workgroups with more chunks in
    // flight move the data faster
    const bool slab = !thin && (p > 64 || q > 64);
    int MB = 0, NB = 0;
    if (thin) {
        nchunk = (int)std::min<int64_t>(1024, std::max<int64_t>(1, (n + kThreads - 1) / kThreads));
    } else if (slab) {
        if (!aligned32(V) || !aligned32(W) || (ldv * es) % 32 || (ldw * es) % 32) {
            latch(170, "gram: operands are not 32-byte aligned");
            return false;
        }
        const int p16 = (p + 15) / 16, q16 = (q + 15) / 16;
        tp = (p16 + 7) / 8;
        tq = (q16 + 7) / 8;
        MB = (p16 + tp - 1) / tp;
        NB = (q16 + tq - 1) / tq;
        const int64_t slab_rows = 128 / (int64_t)es;
        const int64_t max_chunks = std::max<int64_t>(1, (n + slab_rows - 1) / slab_rows);
        // at most kGramSlabWG workgroups: all of them resident at once
        nchunk = (int)std::min<int64_t>({(int64_t)kMaxChunks, max_chunks,
                                         std::max<int64_t>(1, kGramSlabWG / (tp * tq))});
        rpc = (n + nchunk - 1) / nchunk;
        rpc = ((rpc + slab_rows - 1) / slab_rows) * slab_rows;
        nchunk = (int)((n + rpc - 1) / rpc);
    } else {
        if (!aligned32(V) || !aligned32(W) || (ldv * es) % 32 || (ldw * es) % 32) {
            latch(170, "gram: operands are not 32-byte aligned");
            return false;
        }
        g = gram_shape(p, q);
        const int TP = g.GP * g.WP * 16, TQ = g.GQ * g.WQ * 16;
        const int GK = 4 / (g.GP * g.GQ);
        const int R = prec == GCGE_HIP_FP64 ? 4 : 8;
        tp = (p + TP - 1) / TP;
        tq = (q + TQ - 1) / TQ;
        const int64_t unit = (int64_t)4 * R * GK;  // rows one workgroup step covers
        const int64_t max_chunks = std::max<int64_t>(1, (n + unit - 1) / unit);
        nchunk = (int)std::min<int64_t>({(int64_t)kMaxChunks, max_chunks,
                                         std::max<int64_t>(1, (kGramTargetWG + tp * tq - 1) / (tp * tq))});
        rpc = (n + nchunk - 1) / nchunk;
        rpc = ((rpc + unit - 1) / unit) * unit;
        nchunk = (int)((n + rpc - 1) / rpc);
    }
    double *partial = (double *)scratch(sizeof(double) * ((size_t)nchunk * pq + pq));
    if (!partial) return false;
    {
        ProfScope ps(thin ? K_GRAM_THIN : (prec == GCGE_HIP_FP64 ? K_GRAM_MFMA_F64 : K_GRAM_MFMA_F32),
                     (double)n * (p + q) * es + 8.0 * pq, 2.0 * (double)n * pq);
        if (thin) {
            // the kernel wants p <= 4: swap the operands and transpose the output otherwise
            const bool swap = p > 4;
            if (prec == GCGE_HIP_FP64) {
                if (!swap) launch_thin_sized<double>(nchunk, n, V, ldv, p, W, ldw, q, 0, p, partial);
                else launch_thin_sized<double>(nchunk, n, W, ldw, q, V, ldv, p, 1, p, partial);
            } else {
                if (!swap) launch_thin_sized<float>(nchunk, n, V, ldv, p, W, ldw, q, 0, p, partial);
                else launch_thin_sized<float>(nchunk, n, W, ldw, q, V, ldv, p, 1, p, partial);
            }
        } else if (slab) {
            const int64_t blocks = (int64_t)tp * tq * nchunk;
            const int grid = (int)(((blocks + 7) / 8) * 8);
            if (prec == GCGE_HIP_FP64)
                launch_gram_slab<double>(MB, grid, n, V, ldv, p, W, ldw, q, NB, tp, tq, nchunk, rpc,
                                         upper_only ? 1 : 0, partial);
            else
                launch_gram_slab<float>(MB, grid, n, V, ldv, p, W, ldw, q, NB, tp, tq, nchunk, rpc,
                                        upper_only ? 1 : 0, partial);
        } else {
            const int64_t blocks = (int64_t)tp * tq * nchunk;
            const int grid = (int)(((blocks + 7) / 8) * 8);
            if (prec == GCGE_HIP_FP64)
                launch_gram_mfma<double>(g, grid, n, V, ldv, p, W, ldw, q, tp, tq, nchunk, rpc, upper_only ? 1 : 0,
                                         partial);
            else
                launch_gram_mfma<float>(g, grid, n, V, ldv, p, W, ldw, q, tp, tq, nchunk, rpc, upper_only ? 1 : 0,
                                        partial);
        }
    }
    if (!check(hipGetLastError(), "k_gram")) return false;
    double *dG = partial + (int64_t)nchunk * pq;
    // the thin kernel writes every entry, so upper_only only prunes MFMA tiles
    const int up = (!thin && upper_only) ? 1 : 0;
    k_sum_chunks<<<(int)((pq + 63) / 64), kThreads, 0, ctx().stream>>>(partial, nchunk, p, pq, up, dG);
    if (!check(hipGetLastError(), "k_sum_chunks")) return false;
    double *h = host_buf((size_t)pq);
    if (!h) return false;
    if (!check(hipMemcpyAsync(h, dG, sizeof(double) * pq, hipMemcpyDeviceToHost, ctx().stream), "D2H gram"))
        return false;
    if (!sync()) return false;
    for (int j = 0; j < q; j++) std::memcpy(G + (int64_t)j * ldg, h + (int64_t)j * p, sizeof(double) * p);
    return true;
}

bool lincomb(int prec, int64_t n, const void *X, int64_t ldx, int p, void *Y, int64_t ldy, int q, const double *C,
             int ldc, double alpha, double beta)
{
    if (q <= 0 || n <= 0) return true;
    if (failed()) return false;
    if (p <= 0) {
        // empty sum: Y = beta Y
        std::vector<double> b(kMaxCols, beta), a(kMaxCols, 0.0);
        for (int c0 = 0; c0 < q; c0 += kMaxCols)
            launch_axpby(prec, n, Y, ldy, (char *)Y + (size_t)c0 * ldy * elem_size(prec), ldy,
                         std::min(kMaxCols, q - c0), a.data(), b.data(), nullptr);
        return !failed();
    }
    if (p <= 8 && q <= 8) {
        if (((uintptr_t)X & 15u) || ((uintptr_t)Y & 15u) || (ldx * elem_size(prec)) % 16 ||
            (ldy * elem_size(prec)) % 16) {
            latch(171, "lincomb: operands are not 16-byte aligned");
            return false;
        }
        ThinCoef cf;
        for (int j = 0; j < q; j++)
            for (int i = 0; i < p; i++) cf.c[j * p + i] = C[(int64_t)j * ldc + i];
        ProfScope ps(K_LINCOMB_THIN, (double)n * (p + q + (beta != 0.0 ? q : 0)) * elem_size(prec),
                     2.0 * (double)n * p * q);
        if (prec == GCGE_HIP_FP64) launch_lincomb_thin<double>(n, X, ldx, p, Y, ldy, q, cf, alpha, beta);
        else launch_lincomb_thin<float>(n, X, ldx, p, Y, ldy, q, cf, alpha, beta);
        return check(hipGetLastError(), "k_lincomb_thin");
    }
    if (((uintptr_t)Y & 31u) || (ldy * elem_size(prec)) % 32) {
        latch(171, "lincomb: output is not 32-byte aligned");
        return false;
    }
    // column blocks per wavefront pass (RN) and wavefronts per row block
    // (WC): more than 4 column blocks are split over two wavefronts (4 blocks
    // each), which keeps the accumulators small enough for two or more
    // wavefronts per SIMD
    const int nkb = (p + 3) / 4, nqb0 = (q + 15) / 16;
    const bool split = nqb0 > 4;
    const int rn = split ? 4 : nqb0 <= 1 ? 1 : nqb0 <= 2 ? 2 : nqb0 <= 4 ? 4 : 8;
    const int wc = split ? 2 : 1;
    const int nqb = ((nqb0 + rn * wc - 1) / (rn * wc)) * rn * wc;  // zero blocks pad the last groups
    const size_t ncp = (size_t)nkb * nqb * 64;
    const void *dC = nullptr;
    if (prec == GCGE_HIP_FP64) {
        std::vector<double> cp(ncp, 0.0);
        for (int kb = 0; kb < nkb; kb++)
            for (int jb = 0; jb < nqb; jb++)
                for (int l = 0; l < 64; l++) {
                    const int k = 4 * kb + l / 16, j = 16 * jb + l % 16;
                    if (k < p && j < q) cp[((size_t)kb * nqb + jb) * 64 + l] = C[(int64_t)j * ldc + k];
                }
        dC = stage_to_device(cp.data(), sizeof(double) * ncp);
    } else {
        std::vector<float> cp(ncp, 0.0f);
        for (int kb = 0; kb < nkb; kb++)
            for (int jb = 0; jb < nqb; jb++)
                for (int l = 0; l < 64; l++) {
                    const int k = 4 * kb + l / 16, j = 16 * jb + l % 16;
                    if (k < p && j < q) cp[((size_t)kb * nqb + jb) * 64 + l] = (float)C[(int64_t)j * ldc + k];
                }
        dC = stage_to_device(cp.data(), sizeof(float) * ncp);
    }
    if (!dC) return false;
    const int grid = (int)((n + 128 / wc - 1) / (128 / wc));
    {
        ProfScope ps(prec == GCGE_HIP_FP64 ? K_LINCOMB_MFMA_F64 : K_LINCOMB_MFMA_F32, (double)n * (p + q + (beta != 0.0 ? q : 0)) * elem_size(prec),
                     2.0 * (double)n * p * q);
#define GH_LC(T_, RN_) GH_LCW(T_, RN_, 1)
#define GH_LCW(T_, RN_, WC_)                                                                                   \
    k_lincomb_mfma<T_, RN_, WC_><<<grid, kThreads, 0, ctx().stream>>>(n, (const T_ *)X, ldx, p, (T_ *)Y, ldy,  \
                                                                      q, (const T_ *)dC, nkb, nqb, alpha, beta)
        if (prec == GCGE_HIP_FP64) {
            if (wc == 2) GH_LCW(double, 4, 2);
            else if (rn == 1) GH_LC(double, 1);
            else if (rn == 2) GH_LC(double, 2);
            else GH_LC(double, 4);
        } else {
            if (wc == 2) GH_LCW(float, 4, 2);
            else if (rn == 1) GH_LC(float, 1);
            else if (rn == 2) GH_LC(float, 2);
            else GH_LC(float, 4);
        }
#undef GH_LC
#undef GH_LCW
    }
    return check(hipGetLastError(), "k_lincomb");
}

}  // namespace gh
